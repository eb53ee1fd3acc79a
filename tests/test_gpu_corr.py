"""The multiple-tau stress autocorrelator on the device
(src/analysis/correlator.cu:33-110): samples pushed through the C ABI and the
pressure tensors a run pushes by itself every stress_freq steps, against the
FP64 oracle correlator fed the same samples."""
import numpy as np
import pytest

from conftest import KG_LJ, pos4_of, relax_kg

pytestmark = pytest.mark.gpu


def _tiny(fmd, stress_freq=1):
    p4 = pos4_of([[1.0, 1.0, 1.0], [3.0, 3.0, 3.0]])
    return fmd.Desc(p4, 8.0, KG_LJ, stress_freq=stress_freq)


def _same(c, n, c8, n8):
    assert np.array_equal(n, n8)
    scale = max(np.abs(c8).max(), 1e-300)
    assert np.max(np.abs(c - c8)) <= 1e-12 * scale


def test_reference_signals(fmd, engine, orc):
    # tests/test_correlator.cu:6-72
    sid = engine.add(_tiny(fmd))
    const = np.tile([1.0, 0, 0, 0, 0, 0], (64, 1))
    for x in const:
        engine.corr_push(sid, x)
    c, n = engine.corr(sid)
    seen = n[0] > 0
    assert np.allclose(c[0, seen, 0] / n[0, seen], 1.0, atol=1e-5)
    _same(c, n, *orc.corr_run(const))
    engine.corr_reset(sid)
    c, n = engine.corr(sid)
    assert not c.any() and not n.any()
    decay = np.zeros((256, 6))
    decay[:, 0] = np.exp(-np.arange(256, dtype=np.float32) / np.float32(10.0))
    for x in decay:
        engine.corr_push(sid, x)
    c, n = engine.corr(sid)
    assert c[0, 0, 0] / n[0, 0] > c[0, 1, 0] / n[0, 1]
    _same(c, n, *orc.corr_run(decay))
    engine.remove(sid)


def test_pushed_samples_reach_three_levels(fmd, engine, orc):
    sid = engine.add(_tiny(fmd))
    x = np.random.default_rng(3).normal(size=(4200, 6))
    for s in x:
        engine.corr_push(sid, s)
    c, n = engine.corr(sid)
    assert n[2, 0] > 0
    _same(c, n, *orc.corr_run(x))
    engine.remove(sid)


@pytest.mark.parametrize("freq,steps,ens", [(1, 700, "langevin"), (3, 600, "npt")])
def test_run_pushes_pressure_tensor(fmd, engine, orc, freq, steps, ens):
    kw = dict(stress_freq=freq, thermo_freq=freq)
    if ens == "npt":
        kw.update(ensemble=fmd.NPT_NH, T=(1.0, 1.0), Tdamp=0.1, P=(0.5, 0.5), Pdamp=1.0, chain=3, nsteps=steps)
    d = relax_kg(fmd, engine, 40, 50, seed=31, **kw)
    sid = engine.add(d)
    engine.step(sid, steps)
    rows, lost = engine.drain_thermo(sid)
    assert lost == 0 and len(rows) == steps // freq
    assert [r["step"] for r in rows] == list(range(freq, steps + 1, freq))
    samples = np.array([r["press"] for r in rows])
    c, n = engine.corr(sid)
    c8, n8 = orc.corr_run(samples)
    print(f"[corr] {ens} every {freq} steps: {len(samples)} samples, levels used "
          f"{int((n8[:, 0] > 0).sum())}", flush=True)
    _same(c, n, c8, n8)
    engine.remove(sid)


def test_push_without_correlator_is_an_error(fmd, engine):
    sid = engine.add(_tiny(fmd, stress_freq=0))
    with pytest.raises(fmd.FmdError) as ei:
        engine.corr_push(sid, np.zeros(6))
    assert ei.value.code == fmd.api.ERR_ARG
    engine.remove(sid)
