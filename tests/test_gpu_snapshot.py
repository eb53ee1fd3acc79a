"""Exact checkpoint and resume of a system's device state (fmd_snapshot /
fmd_system_resume): a run stopped at step 60, snapshotted, removed and resumed
in another engine continues bit for bit like an uninterrupted run, for
Langevin, Nose-Hoover NVT and NPT, including thermo records, the stress
correlator, the box and the rebuild count.  The reference checkpoints through
LAMMPS data restart files only (src/main.cu:736-752)."""
import numpy as np
import pytest

from conftest import relax_kg

pytestmark = pytest.mark.gpu


def _kw(fmd, ens):
    if ens == "nvt":
        return dict(ensemble=fmd.NVT_NH, T=(1.0, 1.2), Tdamp=0.1, chain=3, nsteps=100)
    if ens == "npt":
        return dict(ensemble=fmd.NPT_NH, T=(1.0, 1.0), Tdamp=0.1, P=(0.0, 1.0), Pdamp=1.0, chain=3, nsteps=100)
    return dict(T=(1.0, 1.1), nsteps=100)


@pytest.mark.parametrize("ens", ["langevin", "nvt", "npt"])
def test_resume_continues_bit_for_bit(fmd, engine, ens):
    d = relax_kg(fmd, engine, 20, 50, seed=51, stress_freq=2, thermo_freq=5, **_kw(fmd, ens))
    a = engine.add(d)
    engine.step(a, 100)
    b = engine.add(d)
    engine.step(b, 60)
    snap = engine.snapshot(b)
    engine.remove(b)
    print(f"[snapshot] {ens}: {len(snap)} bytes for {d.natoms} beads", flush=True)
    e2 = fmd.Engine(0)
    try:
        c = e2.resume(snap)
        assert e2.steps_done(c) == 60
        e2.step(c, 40)
        for x, y in zip(engine.state(a), e2.state(c)):
            assert np.array_equal(x, y)
        assert engine.box(a) == e2.box(c) and engine.state_box(a) == e2.state_box(c)
        assert engine.rebuilds(a) == e2.rebuilds(c)
        assert engine.thermo(a) == e2.thermo(c)
        ra, _ = engine.drain_thermo(a)
        rc, _ = e2.drain_thermo(c)
        assert len(ra) == 20 and ra == rc
        for x, y in zip(engine.corr(a), e2.corr(c)):
            assert np.array_equal(x, y)
        e2.remove(c)
    finally:
        e2.close()
    engine.remove(a)


def test_resume_rejects_damaged_snapshots(fmd, engine):
    d = relax_kg(fmd, engine, 4, 50, seed=52)
    s = engine.add(d)
    snap = engine.snapshot(s)
    engine.remove(s)
    bad_magic = b"X" + snap[1:]
    for bad in (bad_magic, snap[:-8], snap[:100]):
        with pytest.raises(fmd.FmdError) as ei:
            engine.resume(bad)
        assert ei.value.code == fmd.api.ERR_ARG
    # the intact snapshot still resumes
    s2 = engine.resume(snap)
    engine.step(s2, 10)
    engine.remove(s2)
