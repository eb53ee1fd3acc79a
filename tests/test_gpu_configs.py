"""The benchmark configurations at their real sizes, against the oracle:
a 30,000-bead melt (300 x 100) after 1,000 steps; a 1,000,000-bead melt
(10,000 x 100, 32-bit rows) through many rebuilds with Morton re-sorts; a
batch of uneven replicas spanning the three size classes (workload seed
12345, replicas of 2K-64K beads); and a staggered stream of served jobs of
1K-100K beads (one above 65,536, Langevin and Nose-Hoover NVT).  Each checks
forces, potential energy and virial against the FP64 oracle (1e-5), the
exact Verlet pair set of a fresh build, that the list the step path carried
still holds every pair inside rc (the skin trigger), and bit identity of a
replica alone and in a batch or the server."""
import os

import numpy as np
import pytest

from conftest import KG_AP, KG_BP, KG_LJ, rel_err_floor, relax_kg

pytestmark = pytest.mark.gpu
RCS = np.float32(2.5) + np.float32(0.3)


def _keys(ij, n):
    return ij[:, 0].astype(np.int64) * n + ij[:, 1]


def check_against_oracle(fmd, engine, orc, sid, d, label, capsys=None):
    """Forces/PE/virial vs the FP64 oracle, the carried list vs the pairs
    inside rc, then a fresh build vs the exact pair set."""
    p4, _, _ = engine.state(sid)
    n = len(p4)
    L = np.float32(engine.state_box(sid))
    excl = orc.exclusions(n, d.bonds)
    # the list the step path built at its last rebuild still holds every pair
    # inside rc at the current positions
    carried = _keys(engine.pairs(sid), n)
    inside = _keys(orc.pair_set(p4, L, np.float32(2.5), excl, method=1), n)
    pos = np.searchsorted(carried, inside)
    assert np.all(pos < len(carried)) and np.array_equal(carried[np.minimum(pos, len(carried) - 1)], inside), label
    del carried, inside, pos
    f, vir, pe = engine.forces(sid)
    f8, v8 = orc.forces_f64(p4, L, KG_LJ, bonds=d.bonds, btypes=d.bond_types, bp=KG_BP, ang4=d.angles, ap=KG_AP,
                            terms=7)
    floor = np.sqrt(np.mean(np.sum(f8[:, :3] ** 2, axis=1)))
    err_f = rel_err_floor(f[:, :3], f8[:, :3], floor)
    err_pe = abs(pe - f8[:, 3].sum()) / abs(f8[:, 3].sum())
    err_w = np.max(np.abs(np.asarray(vir) - v8)) / np.abs(v8).max()
    del f8
    engine.build_neighbors(sid)
    got = engine.pairs(sid)
    ref = orc.pair_set(p4, L, RCS, excl, method=1)
    same = np.array_equal(got, ref)
    if capsys is not None:
        with capsys.disabled():
            print(f"\n[{label}] {n} beads, {len(ref)} listed pairs, rebuilds {engine.rebuilds(sid)}: "
                  f"force {err_f:.2e}, PE {err_pe:.2e}, virial {err_w:.2e}")
    assert same, label
    assert err_f < 1e-5 and err_pe < 1e-5 and err_w < 1e-5, (label, err_f, err_pe, err_w)
    return len(ref)


def test_small_melt_30k(fmd, engine, orc, capsys):
    """300 x 100 beads (L = 32.80), Langevin T = 1, Tdamp = 1, 1,000 steps."""
    d = relax_kg(fmd, engine, 300, 100, seed=2024, nsteps=1000)
    assert d.natoms == 30000 and abs(float(d.c.L) - 32.802) < 1e-3
    sid = engine.add(d)
    r0 = engine.rebuilds(sid)
    engine.step(sid, 1000)
    assert engine.steps_done(sid) == 1000 and engine.rebuilds(sid) - r0 >= 50
    check_against_oracle(fmd, engine, orc, sid, d, "30k", capsys)
    # the short push-off leaves energy stored in the chains that a bath of
    # Tdamp = 1 has not drained in 1,000 steps of dt = 0.001 (1 time unit), so
    # the kinetic temperature sits above T; the check is that it stays bounded
    t = engine.thermo(sid)
    assert 0.8 < t["temp"] < 1.5
    engine.remove(sid)


def test_single_melt_1m(fmd, engine, orc, capsys):
    """10,000 x 100 beads (L = 105.57), 32-bit rows, 200 steps with a Morton
    re-sort at each of at least 10 rebuilds."""
    d = relax_kg(fmd, engine, 10000, 100, seed=2025)
    assert d.natoms == 1000000
    sid = engine.add(d)
    r0 = engine.rebuilds(sid)
    engine.step(sid, 200)
    assert engine.rebuilds(sid) - r0 >= 10
    check_against_oracle(fmd, engine, orc, sid, d, "1M", capsys)
    engine.remove(sid)


def _relaxed_descs(fmd, engine, specs, ensembles=None):
    from fastmd_amd import workloads
    gen = []
    sids = workloads.relaxed_replicas(engine, specs, descs=gen)
    out = []
    for k, (s, sid, g) in enumerate(zip(specs, sids, gen)):
        p, v, im = engine.state(sid)
        ens = fmd.LANGEVIN if ensembles is None else ensembles[k]
        out.append(workloads.state_desc(g, p, v, im, engine.box(sid), ensemble=ens, T=s["T"],
                                        Tdamp=1.0 if ens == fmd.LANGEVIN else 0.1, seed=s["seed"],
                                        nsteps=100000))
        engine.remove(sid)
    return out


def test_uneven_replicas_three_classes(fmd, engine, orc, capsys):
    """Replicas of the uneven workload (seed 12345) spanning the size classes
    (< 8K, 8K-32K, >= 32K beads, one of 32K-64K): 100 batched steps, every
    replica checked against the oracle and bit-identical alone, and the batch
    bit-identical when it builds the rebuilding replicas' lists beside the
    others' force pass (relaxed apart, they rebuild at different steps)."""
    from fastmd_amd import workloads
    specs = workloads.uneven_replica_specs(256)
    sizes = [s["nchains"] * s["chain_len"] for s in specs]
    pick = []
    for cls in range(3):
        ks = [k for k, b in enumerate(sizes) if fmd.size_class(b) == cls]
        pick += ks[:3]
    assert any(32768 <= sizes[k] <= 65536 for k in pick)
    sub = [specs[k] for k in pick]
    descs = _relaxed_descs(fmd, engine, sub)
    sids = [engine.add(x) for x in descs]
    engine.step(sids, 100)
    states = [engine.state(s) for s in sids]
    for k, (s, x) in enumerate(zip(sids, descs)):
        check_against_oracle(fmd, engine, orc, s, x, f"replica {pick[k]} ({sizes[pick[k]]} beads)",
                             capsys if k % 3 == 0 else None)
    for s in sids:
        engine.remove(s)
    for k in range(len(descs)):
        a = engine.add(descs[k])
        engine.step(a, 100)
        for x, y in zip(engine.state(a), states[k]):
            assert np.array_equal(x, y)
        engine.remove(a)
    os.environ["FMD_STEP_FORK"] = "1"
    try:
        sids = [engine.add(x) for x in descs]
        engine.step(sids, 100)
        for s, ref in zip(sids, states):
            for x, y in zip(engine.state(s), ref):
                assert np.array_equal(x, y)
        for s in sids:
            engine.remove(s)
    finally:
        del os.environ["FMD_STEP_FORK"]


def test_served_job_stream(fmd, engine, orc):
    """Twelve jobs of 1K-100K beads (log-uniform, one above 65,536: 32-bit
    rows), 30% Nose-Hoover NVT, T in [0.8, 1.2], arriving over several
    launches: every job's final state equals its alone run bit for bit."""
    import time

    rng = np.random.default_rng(777)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(100000), 11)).tolist() + [70000]
    specs = []
    for k, b in enumerate(sizes):
        clen = int(round(np.exp(rng.uniform(np.log(20), np.log(200)))))
        specs.append({"nchains": max(1, int(round(b / clen))), "chain_len": clen,
                      "T": float(rng.uniform(0.8, 1.2)), "seed": 3000 + k, "gen_seed": 4000 + k})
    ens = [fmd.NVT_NH if rng.random() < 0.3 else fmd.LANGEVIN for _ in specs]
    ens[0] = fmd.NVT_NH
    descs = _relaxed_descs(fmd, engine, specs, ens)
    assert max(x.natoms for x in descs) > 65536
    lens = [int(x) for x in rng.integers(60, 400, len(descs))]
    alone = []
    for x, n in zip(descs, lens):
        s = engine.add(x)
        engine.step(s, n)
        alone.append(engine.state(s))
        engine.remove(s)
    srv = fmd.Server(0, graph_steps=20)
    jobs = []
    for x, n in zip(descs, lens):
        jobs.append(srv.submit(x, n))
        time.sleep(0.03)
    srv.drain()
    for j, st, n, x in zip(jobs, alone, lens, descs):
        r = srv.result(j, state=True)
        assert r["status"] == 0 and r["steps"] == n
        assert np.array_equal(r["pos4"], st[0]) and np.array_equal(r["vel"], st[1]), (x.natoms, n)
    st = srv.stats()
    assert st["graph_launches"] >= max(lens) // 20
    srv.close()
