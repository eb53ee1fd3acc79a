"""Bit-exact parity of the device Morton keys, cell ids and Verlet pair sets
with the oracle (src/core/morton.cu:12-28, src/neighbor/verlet_list.cu:58-335)."""
import numpy as np
import pytest

from conftest import KG_LJ, pos4_of, relax_kg

pytestmark = pytest.mark.gpu


def _as_set(ij):
    return set(map(tuple, np.asarray(ij).tolist()))


def test_morton_keys_and_cells_reference_input(fmd, engine, orc, golden):
    # tests/test_morton.cu:12-20 input (closed-form positions, L=10)
    xyz = golden["morton64_xyz"]
    L = np.float32(10.0)
    p4 = pos4_of(xyz)
    sid = engine.add(fmd.Desc(p4, L, KG_LJ))
    keys = engine.morton_keys(sid)
    ref = np.array([orc.morton3d(*x, L) for x in xyz], np.uint32)
    assert np.array_equal(keys, ref)
    # the reference's permutation: stable sort by (key, index)
    assert np.array_equal(np.argsort(keys, kind="stable"), orc.morton_perm(p4, L))
    cells, nx = engine.cell_ids(sid)
    rcells, rnx = orc.cell_ids(p4, L, np.float32(2.5) + np.float32(0.3))
    assert nx == rnx and np.array_equal(cells, rcells)
    engine.remove(sid)


@pytest.mark.parametrize("n,L,seed", [(4000, 16.77, 1), (3000, 7.5, 2), (500, 8.0, 3)])
def test_morton_keys_and_cells_random(fmd, engine, orc, n, L, seed):
    rng = np.random.default_rng(seed)
    L = np.float32(L)
    xyz = (rng.random((n, 3)) * L).astype(np.float32)
    xyz[:8] = [[0, 0, 0], [np.nextafter(L, 0, dtype=np.float32)] * 3, [L / 2] * 3, [L / 3] * 3,
               [1e-7, L - 1e-5, 0.5], [2.8, 5.6, 8.4], [0.0, 2.8, 0.0], [L / 4, L / 5, L / 7]]
    p4 = pos4_of(xyz)
    sid = engine.add(fmd.Desc(p4, L, KG_LJ, force_cap=10.0))
    keys = engine.morton_keys(sid)
    assert np.array_equal(keys, np.array([orc.morton3d(*x, L) for x in xyz], np.uint32))
    cells, nx = engine.cell_ids(sid)
    rcells, rnx = orc.cell_ids(p4, L, np.float32(2.8))
    assert nx == rnx and np.array_equal(cells, rcells)
    engine.remove(sid)


def test_pairs_reference_lj64(fmd, engine, orc, golden):
    # tests/test_lj.cu:8-27: N=64, L=8, rc=2.5, skin=0.5 -> 2 cells per axis
    # (the de-duplicated small-box stencil)
    p4 = pos4_of(golden["lj64_xyz"])
    sid = engine.add(fmd.Desc(p4, 8.0, KG_LJ, rc=2.5, skin=0.5))
    got = engine.pairs(sid)
    ref = orc.pair_set(p4, np.float32(8.0), np.float32(3.0))
    assert len(ref) > 0
    assert np.array_equal(got, ref)
    engine.remove(sid)


# (300, 17.0): sparse cells, so a tile's candidate cells exceed the LDS
# staging limits and the list build takes its global-memory scan
@pytest.mark.parametrize("n,L,seed", [(4000, 16.77, 4), (2000, 6.0, 5), (700, 9.3, 6), (300, 17.0, 7)])
def test_pairs_random(fmd, engine, orc, n, L, seed):
    rng = np.random.default_rng(seed)
    p4 = pos4_of((rng.random((n, 3)) * L).astype(np.float32))
    sid = engine.add(fmd.Desc(p4, L, KG_LJ, force_cap=10.0))
    got = engine.pairs(sid)
    ref = orc.pair_set(p4, np.float32(L), np.float32(2.5) + np.float32(0.3))
    assert np.array_equal(got, ref)
    engine.remove(sid)


def test_pairs_kg_melt_with_exclusions_after_steps(fmd, engine, orc):
    d = relax_kg(fmd, engine, 10, 100, seed=7)
    sid = engine.add(d)
    engine.step(sid, 60)  # several rebuilds and Morton re-sorts
    engine.build_neighbors(sid)
    p4, _, _ = engine.state(sid)
    excl = orc.exclusions(len(p4), d.bonds)
    ref = orc.pair_set(p4, np.float32(d.c.L), np.float32(2.5) + np.float32(0.3), excl)
    got = engine.pairs(sid)
    assert np.array_equal(got, ref)
    # bonded partners never appear in the list
    bset = _as_set(d.bonds) | _as_set(d.bonds[:, ::-1])
    assert not (_as_set(got) & bset)
    engine.remove(sid)


def test_pairs_and_forces_with_32bit_rows(fmd, engine, orc, monkeypatch):
    """Replicas above 65536 slots keep 32-bit neighbor rows; FMD_IDX32 selects
    that layout for a small melt, whose pair set and forces must not change."""
    d = relax_kg(fmd, engine, 10, 100, seed=9)
    sids = {}
    for wide in (False, True):
        if wide:
            monkeypatch.setenv("FMD_IDX32", "1")
        else:
            monkeypatch.delenv("FMD_IDX32", raising=False)
        sids[wide] = engine.add(d)
    monkeypatch.delenv("FMD_IDX32", raising=False)
    p4, _, _ = engine.state(sids[False])
    excl = orc.exclusions(len(p4), d.bonds)
    ref = orc.pair_set(p4, np.float32(d.c.L), np.float32(2.5) + np.float32(0.3), excl)
    f16 = engine.forces(sids[False])
    f32 = engine.forces(sids[True])
    assert np.array_equal(engine.pairs(sids[False]), ref)
    assert np.array_equal(engine.pairs(sids[True]), ref)
    for a, b in zip(f16, f32):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    # both layouts place the same rows, so 60 steps (several rebuilds through
    # the step path) give the same bits, and the lists stay the exact pair set
    engine.step([sids[False], sids[True]], 60)
    assert engine.rebuilds(sids[False]) >= 3
    for a, b in zip(engine.state(sids[False]), engine.state(sids[True])):
        assert np.array_equal(a, b)
    p4, _, _ = engine.state(sids[True])
    engine.build_neighbors(sids[True])
    ref = orc.pair_set(p4, np.float32(engine.state_box(sids[True])), np.float32(2.5) + np.float32(0.3), excl)
    assert np.array_equal(engine.pairs(sids[True]), ref)
    for s in sids.values():
        engine.remove(s)
