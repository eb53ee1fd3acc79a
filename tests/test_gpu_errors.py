"""Failure signals (a replica that stops on its own, the others unchanged):
a FENE bond stretched past R0 (the reference clamps, src/force/fene.cu:33),
a non-finite velocity, and a neighbor count above the list capacity (the
reference truncates, src/neighbor/verlet_list.cu:281-285), each in one batch
with healthy melts.  Every failing replica stops with its code at the step
the failure happened, alone and in the batch alike, and the healthy
replicas equal their alone runs bit for bit."""
import numpy as np
import pytest

from conftest import KG_AP, KG_BP, KG_LJ, pos4_of, relax_kg

pytestmark = pytest.mark.gpu
F = np.float32


def _chain_desc(fmd, vel):
    xyz = np.stack([5.0 + 0.97 * np.arange(10), np.full(10, 5.0), np.full(10, 5.0)], 1)
    bonds = np.array([[i, i + 1] for i in range(9)], np.int32)
    ang = np.array([[i, i + 1, i + 2, 0] for i in range(8)], np.int32)
    return fmd.Desc(pos4_of(xyz), 20.0, KG_LJ, vel=vel, bonds=bonds, bond_params=KG_BP, angles=ang,
                    angle_params=KG_AP, Tdamp=1e10)


def stretch_desc(fmd):
    """Beads 0 and 1 of a straight chain fly apart at 60 sigma/tau each."""
    v = np.zeros((10, 3), F)
    v[0, 0], v[1, 0] = -60.0, 60.0
    return _chain_desc(fmd, v)


def nan_desc(fmd):
    v = np.zeros((10, 3), F)
    v[3, 1] = np.nan
    return _chain_desc(fmd, v)


def overflow_desc(fmd):
    """Two blocks of 150 beads (5 x 5 x 6 at spacing 0.35, LJ sigma 0.1) fly
    into each other along x, offset so no two beads meet: each bead starts
    with 149 neighbors within rc + skin (capacity 256 at this mean density)
    and ends with nearly 300."""
    g = np.arange(150)
    blk = np.stack([g % 5, (g // 5) % 5, g // 25], 1) * 0.35
    xyz = np.concatenate([blk + [5.0, 5.0, 5.0], blk + [10.0, 5.175, 5.175]])
    v = np.zeros((300, 3), F)
    v[:150, 0], v[150:, 0] = 10.0, -10.0
    return fmd.Desc(pos4_of(xyz), 20.0, np.array([1.0, 0.1], F), vel=v, Tdamp=1e10, dt=0.002)


def test_failures_stop_alone_and_leave_others_unchanged(fmd, engine, orc):
    healthy = [relax_kg(fmd, engine, 8, 50, seed=61), relax_kg(fmd, engine, 20, 40, seed=62)]
    makers = {fmd.ERR_STRETCH: stretch_desc, fmd.ERR_NAN: nan_desc, fmd.ERR_NBR_OVERFLOW: overflow_desc}
    nst = 300
    # alone
    alone_h = []
    for d in healthy:
        s = engine.add(d)
        engine.step(s, nst)
        alone_h.append((engine.state(s), engine.thermo(s)))
        engine.remove(s)
    alone_f = {}
    for code, mk in makers.items():
        s = engine.add(mk(fmd))
        with pytest.raises(fmd.FmdError) as ex:
            engine.step(s, nst)
        assert ex.value.code == code
        assert engine.error(s) == code
        alone_f[code] = (engine.steps_done(s), engine.state(s))
        engine.remove(s)
    # the same replicas in one batch
    ids = [engine.add(healthy[0])] + [engine.add(mk(fmd)) for mk in makers.values()] + [engine.add(healthy[1])]
    with pytest.raises(fmd.FmdError) as ex:
        engine.step(ids, nst)
    assert ex.value.code in makers
    for (code, _), s in zip(makers.items(), ids[1:4]):
        assert engine.error(s) == code
        step, st = alone_f[code]
        assert engine.steps_done(s) == step
        for x, y in zip(engine.state(s), st):
            assert np.array_equal(x, y, equal_nan=True)
        # stopped: a further run leaves it where it is
        with pytest.raises(fmd.FmdError):
            engine.step(s, 5)
        assert engine.steps_done(s) == step
    for s, (st, th) in zip([ids[0], ids[4]], alone_h):
        assert engine.error(s) == fmd.OK and engine.steps_done(s) == nst
        for x, y in zip(engine.state(s), st):
            assert np.array_equal(x, y)
        assert engine.thermo(s) == th
    for s in ids:
        engine.remove(s)

    # what each stop state shows
    step, (p, v, im) = alone_f[fmd.ERR_NAN]
    assert step == 0 and not np.all(np.isfinite(p[3, :3]))
    step, (p, v, im) = alone_f[fmd.ERR_STRETCH]
    r = np.linalg.norm(p[1, :3] - p[0, :3])
    assert 1 <= step < 20 and r >= 1.5
    s = engine.add(stretch_desc(fmd))  # one step earlier the bond was still inside R0
    if step > 1:
        engine.step(s, step - 1)
        p1, _, _ = engine.state(s)
        assert np.linalg.norm(p1[1, :3] - p1[0, :3]) < 1.5
    engine.remove(s)
    step, (p, v, im) = alone_f[fmd.ERR_NBR_OVERFLOW]
    assert 20 < step < nst
    ij = orc.pair_set(p, 20.0, 2.8)
    assert np.max(np.bincount(ij[:, 0], minlength=300)) > 256
