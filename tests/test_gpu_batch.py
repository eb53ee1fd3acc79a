"""Batched multi-replica stepping: each replica's result is bit-identical
alone and in the mix, with the plain one-system path and through the server."""
import os

import numpy as np
import pytest

from conftest import KG_AP, KG_BP, KG_LJ, relax_kg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trio(fmd, engine):
    out = []
    for k, (nch, clen, ens) in enumerate([(10, 100, 0), (12, 50, 1), (4, 200, 2)]):
        m = relax_kg(fmd, engine, nch, clen, seed=31 + k)
        out.append(fmd.Desc(m.pos4, m.c.L, KG_LJ, vel=m.vel, image=m.image, bonds=m.bonds, bond_params=KG_BP,
                            angles=m.angles, angle_params=KG_AP, ensemble=ens, T=(1.0, 1.1), Tdamp=0.5,
                            P=(0.5, 1.0), seed=500 + k, nsteps=300, thermo_freq=25))
    return out


def _run_alone(fmd, d, nsteps, plain=False):
    e = fmd.Engine(0)
    sid = e.add(d)
    if plain:
        e.step_plain(sid, nsteps)
    else:
        e.step(sid, nsteps)
    st = e.state(sid)
    th = e.drain_thermo(sid)[0]
    e.close()
    return st, th


def test_alone_equals_batched(fmd, trio):
    alone = [_run_alone(fmd, d, 137) for d in trio]
    e = fmd.Engine(0)
    sids = [e.add(d) for d in trio]
    e.step(sids, 137)
    for sid, (st, th) in zip(sids, alone):
        p, v, im = e.state(sid)
        assert np.array_equal(p, st[0]) and np.array_equal(v, st[1]) and np.array_equal(im, st[2])
        assert e.drain_thermo(sid)[0] == th
    e.close()


def test_forked_step_equals_alone(fmd, trio):
    """With FMD_STEP_FORK=1 the replicas that rebuild in a step sort, build their
    lists and compute their forces on a second stream beside the other
    replicas' force pass: Langevin, NVT-NH and NPT-NH replicas still equal
    their alone runs (one-stream step) bit for bit, thermo rows included."""
    alone = [_run_alone(fmd, d, 137) for d in trio]
    os.environ["FMD_STEP_FORK"] = "1"
    try:
        e = fmd.Engine(0)
        sids = [e.add(d) for d in trio]
        e.step(sids, 137)
        for sid, (st, th) in zip(sids, alone):
            p, v, im = e.state(sid)
            assert np.array_equal(p, st[0]) and np.array_equal(v, st[1]) and np.array_equal(im, st[2])
            assert e.drain_thermo(sid)[0] == th
        e.close()
    finally:
        del os.environ["FMD_STEP_FORK"]


def test_plain_equals_graph(fmd, trio):
    a, tha = _run_alone(fmd, trio[0], 77, plain=True)
    b, thb = _run_alone(fmd, trio[0], 77, plain=False)
    assert all(np.array_equal(x, y) for x, y in zip(a, b)) and tha == thb


def test_split_runs_equal_one_run(fmd, trio):
    """Stepping 60 + 40 equals stepping 100 (device state carries across calls)."""
    e = fmd.Engine(0)
    s1 = e.add(trio[1])
    s2 = e.add(trio[1])
    e.step(s1, 60)
    e.step(s1, 40)
    e.step(s2, 100)
    assert all(np.array_equal(x, y) for x, y in zip(e.state(s1), e.state(s2)))
    e.close()


def test_server_results_equal_alone(fmd, trio):
    lens = [120, 45, 200]
    alone = [_run_alone(fmd, d, n)[0] for d, n in zip(trio, lens)]
    srv = fmd.Server(0, graph_steps=20)
    # block shapes as an autotuned engine reports them: any force shape gives
    # the same results; the list build has one shape
    srv.set_blocks({(0, 0): 64, (0, 1): 128, (0, 2): 256, (1, 0): 256, (1, 1): 256, (1, 2): 256})
    with pytest.raises(fmd.FmdError):
        srv.set_blocks({(1, 0): 64})
    jobs = [srv.submit(d, n) for d, n in zip(trio, lens)]
    srv.drain()
    for j, st, n in zip(jobs, alone, lens):
        r = srv.result(j, state=True)
        assert r["status"] == 0 and r["steps"] == n
        assert np.array_equal(r["pos4"], st[0]) and np.array_equal(r["vel"], st[1])
    srv.close()


def test_server_budget_queues_jobs(fmd, trio):
    """A budget that holds one job at a time: jobs wait, none is dropped, FIFO order."""
    b = max(fmd.system_bytes(d) for d in trio)
    srv = fmd.Server(0, hbm_budget=int(b * 1.2), graph_steps=20)
    jobs = [srv.submit(d, 60) for d in trio]
    srv.drain()
    res = [srv.result(j) for j in jobs]
    assert all(r["status"] == 0 for r in res)
    for a, c in zip(res, res[1:]):
        assert c["t_admit"] >= a["t_done"] - 1e-6
    srv.close()


def test_server_tight_budget_many_small_jobs(fmd, engine):
    """Admission and the engine charge a job the same bytes: with a budget of
    three 1,000-bead jobs, twelve such jobs all run (none is refused by the
    engine after admission), at most three at a time, and each equals its
    alone run."""
    d = relax_kg(fmd, engine, 10, 100, seed=71)
    descs = []
    for k in range(12):
        x = fmd.Desc(d.pos4, d.c.L, KG_LJ, vel=d.vel, image=d.image, bonds=d.bonds, bond_params=KG_BP,
                     angles=d.angles, angle_params=KG_AP, T=(1.0, 1.0), seed=900 + k, nsteps=1000)
        descs.append(x)
    b = fmd.system_bytes(descs[0])
    assert b % (2 << 20) == 0
    e = fmd.Engine(0, hbm_budget=3 * b)
    sids = [e.add(x) for x in descs[:3]]
    assert e.bytes_in_use() == 3 * b
    with pytest.raises(fmd.FmdError) as ex:
        e.add(descs[3])
    assert ex.value.code == fmd.ERR_BUDGET
    e.remove(sids[0])
    sids[0] = e.add(descs[3])  # a block from the cache is charged the same
    assert e.bytes_in_use() == 3 * b
    e.close()
    alone = [_run_alone(fmd, x, 60 + 10 * k)[0] for k, x in enumerate(descs)]
    srv = fmd.Server(0, hbm_budget=3 * b, graph_steps=20)
    jobs = [srv.submit(x, 60 + 10 * k) for k, x in enumerate(descs)]
    srv.drain()
    res = [srv.result(j, state=True) for j in jobs]
    assert all(r["status"] == 0 for r in res), [r["status"] for r in res]
    for r, st in zip(res, alone):
        assert np.array_equal(r["pos4"], st[0]) and np.array_equal(r["vel"], st[1])
    # never more than three resident: each admission waits for a completion
    ev = sorted([(r["t_admit"], 1) for r in res] + [(r["t_done"], -1) for r in res])
    live = 0
    for _, x in ev:
        live += x
        assert live <= 3
    srv.close()


def test_autotune_per_class_is_bit_identical(fmd, engine):
    """A batch spanning the three size classes (4,000, 10,000 and 35,000
    beads).  Autotuning times its candidates on copies of the systems, so the
    systems are untouched, and picks per class a force workgroup shape and a
    neighbor layout.  Any setting -- the tuned one, or shapes of 64 / 128
    threads with direct (unstaged) neighbor gathers set before creation or
    mid-run -- gives the same bits as the defaults over 100 steps with
    rebuilds."""
    descs = [relax_kg(fmd, engine, n, 100, seed=80 + k, T=(1.0, 1.0), nsteps=1000)
             for k, n in enumerate((40, 100, 350))]
    assert [fmd.size_class(d.natoms) for d in descs] == [0, 1, 2]

    def run(setup=None, tune=False, mid=None):
        e = fmd.Engine(0)
        if setup:
            setup(e)
        sids = [e.add(d) for d in descs]
        info = None
        if tune:
            before = [e.state(s) for s in sids]
            tuned = e.autotune(sids, reps=2)
            for s, b in zip(sids, before):
                assert all(np.array_equal(x, y) for x, y in zip(e.state(s), b))
            info = (tuned, e.autotune_times())
        e.step(sids, 40)
        if mid:
            mid(e)
        e.step(sids, 60)
        out = [e.state(s) for s in sids]
        e.close()
        return out, info

    ref, _ = run()
    got, (tuned, times) = run(tune=True)
    assert sorted(times) == [0, 1, 2] and all(len(times[c]) == 6 for c in times)
    assert all(tuned[(0, c)] in (64, 128, 256) and tuned[(2, c)] in (0, 1) for c in range(3))

    def shapes(e):
        for c, b in zip(range(3), (64, 128, 64)):
            e.set_tuned(0, c, b)
            e.set_tuned(2, c, 0)

    def mid(e):
        for c in range(3):
            e.set_tuned(2, c, 1 - e.tuned()[(2, c)])

    for variant in (got, run(setup=shapes)[0], run(mid=mid)[0]):
        for a, b in zip(variant, ref):
            assert all(np.array_equal(x, y) for x, y in zip(a, b))
