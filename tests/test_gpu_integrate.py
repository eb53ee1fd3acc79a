"""Integrator parity: BAOAB Langevin (src/integrate/langevin.cu:34-130),
Nose-Hoover NVT and NPT (src/integrate/nose_hoover.cu) against the oracle's MD
driver run in the reference's operator order with the same Philox noise, the
NVE limit of tests/test_integrator.cu, thermo closed forms of
tests/test_thermo.cu and tests/test_mass.cu, and temperature/pressure control."""
import numpy as np
import pytest

from conftest import KG_AP, KG_BP, KG_LJ, pos4_of, relax_kg

pytestmark = pytest.mark.gpu


def _oracle_sys(d, ensemble, T=(1.0, 1.0), Tdamp=1.0, P=(0.0, 0.0), Pdamp=1.0, nsteps_total=1000, seed=42):
    return {"pos4": d.pos4.copy(), "vel3": d.vel.copy(), "img3": d.image.copy(), "masses": np.ones(16, np.float32),
            "lj": KG_LJ, "bonds": d.bonds, "btype": d.bond_types, "bp": KG_BP, "ang4": d.angles, "ap": KG_AP,
            "L": float(d.c.L), "ensemble": ensemble, "T_start": T[0], "T_stop": T[1], "Tdamp": Tdamp,
            "P_start": P[0], "P_stop": P[1], "Pdamp": Pdamp, "chain": 3, "seed": seed,
            "nsteps_total": nsteps_total}


def _min_image(d, L):
    return d - L * np.rint(d / L)


@pytest.fixture(scope="module")
def melt(fmd, engine):
    return relax_kg(fmd, engine, 8, 100, seed=21)


@pytest.mark.parametrize("ensemble", [0, 1, 2])
def test_trajectory_matches_oracle(fmd, engine, orc, melt, ensemble):
    """40 steps from the same state: positions, velocities, box and per-step
    thermo of the device agree with the oracle driver."""
    kw = dict(ensemble=ensemble, T=(1.0, 1.2), Tdamp=(1.0 if ensemble == 0 else 0.1), P=(0.0, 1.0), Pdamp=1.0,
              nsteps=1000, seed=99, thermo_freq=1)
    d = fmd.Desc(melt.pos4, melt.c.L, KG_LJ, vel=melt.vel, image=melt.image, bonds=melt.bonds,
                 bond_params=KG_BP, angles=melt.angles, angle_params=KG_AP, chain=3, **kw)
    sid = engine.add(d)
    nst = 40
    engine.step(sid, nst)
    p, v, im = engine.state(sid)
    rec, lost = engine.drain_thermo(sid)
    Lg = engine.box(sid)
    engine.remove(sid)
    o = _oracle_sys(melt, ensemble, T=kw["T"], Tdamp=kw["Tdamp"], P=kw["P"], seed=99)
    po, vo, io, Lo, th = orc.md_run(o, nst)
    assert abs(Lg - Lo) <= 1e-6 * Lo
    dx = _min_image(p[:, :3].astype(np.float64) - po[:, :3], Lo)
    assert np.max(np.abs(dx)) < 2e-4
    assert np.max(np.abs(v - vo)) < 5e-3
    assert lost == 0 and len(rec) == nst
    for k, r in enumerate(rec):
        t = th[k]
        assert r["step"] == k + 1
        assert abs(r["temp"] - t[3]) <= 2e-4 * t[3]
        assert abs(r["pe"] - t[2]) <= 2e-4 * abs(t[2])
        assert np.max(np.abs(np.array(r["press"]) - t[4:10])) <= 2e-3 * max(1.0, abs(t[4]))


def test_nve_limit_reference_input(fmd, engine, golden):
    # tests/test_integrator.cu:9-127: 64-atom lattice, L=8, Tdamp=1e10, 1000 steps, drift < 1%
    d = fmd.Desc(pos4_of(golden["nve64_xyz"]), 8.0, KG_LJ, vel=golden["nve64_vel"], rc=2.5, skin=0.5,
                 Tdamp=1e10, seed=12345)
    sid = engine.add(d)
    t0 = engine.thermo(sid)
    e0 = (t0["ke"] + t0["pe"]) * 64
    engine.step(sid, 1000)
    t1 = engine.thermo(sid)
    e1 = (t1["ke"] + t1["pe"]) * 64
    engine.remove(sid)
    assert abs(e1 - e0) / abs(e0) < 0.01


def test_nve_drift_kg_melt(fmd, engine, melt):
    """Energy conservation of the whole KG force field in the NVE limit."""
    d = fmd.Desc(melt.pos4, melt.c.L, KG_LJ, vel=melt.vel, image=melt.image, bonds=melt.bonds,
                 bond_params=KG_BP, angles=melt.angles, angle_params=KG_AP, Tdamp=1e10, thermo_freq=100)
    sid = engine.add(d)
    t0 = engine.thermo(sid)
    e0 = t0["ke"] + t0["pe"]
    engine.step(sid, 2000)
    rec, _ = engine.drain_thermo(sid)
    engine.remove(sid)
    e = np.array([r["ke"] + r["pe"] for r in rec])
    # per-atom energy drift over 2 tau (dt = 0.001) well below the thermal scale
    assert np.max(np.abs(e - e0)) < 2e-3 * abs(e0) + 2e-3


def test_thermo_closed_forms(fmd, engine):
    # tests/test_thermo.cu:6-44: unit v_x -> KE/N = 0.5, T = 1/3 (atoms too far apart to interact)
    g = np.arange(96)
    xyz = np.stack([(g % 4) * 4.0, ((g // 4) % 4) * 4.0, (g // 16) * 4.0], 1) + 0.5
    d = fmd.Desc(pos4_of(xyz), 24.0, KG_LJ, vel=np.tile([1.0, 0.0, 0.0], (96, 1)))
    sid = engine.add(d)
    t = engine.thermo(sid)
    engine.remove(sid)
    assert abs(t["ke"] - 0.5) < 1e-12 and abs(t["temp"] - 1.0 / 3.0) < 1e-12
    # tests/test_mass.cu:66-144: masses 1 and 2 -> sum m vx^2 = 3
    xyz2 = np.array([[1.0, 1.0, 1.0], [10.0, 10.0, 10.0]])
    d2 = fmd.Desc(pos4_of(xyz2, [0, 1]), 20.0, np.ones(8, np.float32), ntypes=2, vel=[[1, 0, 0], [1, 0, 0]],
                  masses=[1.0, 2.0])
    sid = engine.add(d2)
    t = engine.thermo(sid)
    assert abs(t["press"][0] * 20.0 ** 3 - 3.0) < 1e-9
    assert abs(t["ke"] * 2 - 1.5) < 1e-12
    # half-kick mass scaling (tests/test_mass.cu:21-63): dv = dt/2 F / m
    engine.remove(sid)


@pytest.mark.parametrize("ensemble,T", [(0, 1.2), (1, 1.5)])
def test_temperature_control(fmd, engine, melt, ensemble, T):
    # tests/test_nose_hoover.cu:195-340: <T> within 0.1 of the target
    d = fmd.Desc(melt.pos4, melt.c.L, KG_LJ, vel=melt.vel, image=melt.image, bonds=melt.bonds,
                 bond_params=KG_BP, angles=melt.angles, angle_params=KG_AP, ensemble=ensemble,
                 T=(T, T), Tdamp=(1.0 if ensemble == 0 else 0.5), thermo_freq=10, nsteps=6000)
    sid = engine.add(d)
    engine.step(sid, 6000)
    rec, _ = engine.drain_thermo(sid)
    engine.remove(sid)
    temps = np.array([r["temp"] for r in rec[len(rec) // 3:]])
    assert abs(temps.mean() - T) < 0.1


def test_npt_pressure_control(fmd, engine, melt):
    """NPT-NH drives <P> to the target (scalar P = trace/3)."""
    d = fmd.Desc(melt.pos4, melt.c.L, KG_LJ, vel=melt.vel, image=melt.image, bonds=melt.bonds,
                 bond_params=KG_BP, angles=melt.angles, angle_params=KG_AP, ensemble=2, T=(1.0, 1.0),
                 Tdamp=0.1, P=(1.0, 1.0), Pdamp=1.0, thermo_freq=10, nsteps=8000)
    sid = engine.add(d)
    engine.step(sid, 8000)
    rec, _ = engine.drain_thermo(sid)
    engine.remove(sid)
    late = rec[len(rec) // 2:]
    P = np.array([(r["press"][0] + r["press"][3] + r["press"][5]) / 3.0 for r in late])
    T = np.array([r["temp"] for r in late])
    assert abs(P.mean() - 1.0) < 0.25
    assert abs(T.mean() - 1.0) < 0.05


def test_rg_reference_cases_and_melt(fmd, engine, orc):
    """Chain-averaged Rg on the device (src/analysis/rg.cu) against the
    reference's own cases (tests/test_rg.cu: 2.23607 and 1.63299) and against
    the FP64 oracle on a KG melt after Morton re-sorts and box wraps."""
    p = np.zeros((8, 4), np.float32)
    p[:, 0] = [1, 3, 5, 7, 2, 4, 6, 8]
    sid = engine.add(fmd.Desc(pos4_of(p[:, :3]), 10.0, KG_LJ, mol=[0, 0, 0, 0, 1, 1, 1, 1], force_cap=10.0))
    assert abs(engine.rg(sid) - 2.23607) < 1e-5
    engine.remove(sid)
    p = np.zeros((3, 4), np.float32)
    p[:, 0] = [9, 1, 3]
    im = np.zeros((3, 3), np.int32)
    im[1:, 0] = 1
    sid = engine.add(fmd.Desc(pos4_of(p[:, :3]), 10.0, KG_LJ, image=im, mol=[0, 0, 0], force_cap=10.0))
    assert abs(engine.rg(sid) - 1.63299) < 1e-5
    engine.remove(sid)
    d = relax_kg(fmd, engine, 20, 50, seed=31)
    sid = engine.add(d)
    engine.step(sid, 100)
    p4, _, im = engine.state(sid)
    rg, per = engine.rg(sid, per_chain=True)
    assert len(per) == 20
    ref = orc.rg(p4, im, d.mol, np.float32(engine.state_box(sid)))
    assert abs(rg - ref) <= 1e-6 * ref
    assert 2.0 < rg < 6.0  # ~ sqrt(N/6) b for N = 50 beads
    engine.remove(sid)


@pytest.mark.parametrize("cap", [0.0, 50.0])
def test_one_run_equals_single_steps(fmd, engine, cap):
    """Within a Langevin run the force pass also gives the next step's
    pre-force half kick; the last step of a run does not, and the next run's
    pre-force kernel gives it from the stored force.  Both routes give the
    same bits: one 30-step run equals 30 one-step runs, including under the
    push-off force cap (with its speed limit)."""
    d = relax_kg(fmd, engine, 12, 40, seed=77, T=(1.0, 1.2), nsteps=100)
    a = engine.add(d)
    b = engine.add(d)
    if cap > 0.0:
        engine.set_force_cap(a, cap)
        engine.set_force_cap(b, cap)
    engine.step(a, 30)
    for _ in range(30):
        engine.step(b, 1)
    for x, y in zip(engine.state(a), engine.state(b)):
        assert np.array_equal(x, y)
    assert engine.thermo(a) == engine.thermo(b)
    engine.remove(a)
    engine.remove(b)


def test_thermo_after_unobserved_steps(fmd, engine):
    """Langevin steps that record nothing sum no FP64 partials; the last step
    of a run always does, so the thermo read after a run with no thermo
    output is that of the final state: its kinetic energy and kinetic stress
    equal sums over the returned velocities, and its step is the last one."""
    d = relax_kg(fmd, engine, 12, 40, seed=78, T=(1.0, 1.2), nsteps=100, thermo_freq=0)
    sid = engine.add(d)
    s0 = engine.steps_done(sid)
    engine.step(sid, 37)
    t = engine.thermo(sid)
    _, v, _ = engine.state(sid)
    vd = v.astype(np.float64)
    n = len(vd)
    assert t["step"] == s0 + 37
    assert abs(t["ke"] - 0.5 * np.sum(vd * vd) / n) < 1e-9 * max(1.0, t["ke"])
    # PE and pressure tensor: the final step's virial equals a fresh evaluation
    # at the returned state, plus the kinetic tensor of the returned velocities
    _, vir, pe = engine.forces(sid)
    V = float(engine.box(sid)) ** 3
    kin = np.array([vd[:, 0] @ vd[:, 0], vd[:, 0] @ vd[:, 1], vd[:, 0] @ vd[:, 2],
                    vd[:, 1] @ vd[:, 1], vd[:, 1] @ vd[:, 2], vd[:, 2] @ vd[:, 2]])
    assert abs(t["pe"] - pe / n) <= 1e-12 * abs(pe / n)
    scale = np.max(np.abs(kin) + np.abs(np.asarray(vir))) / V
    assert np.max(np.abs(np.asarray(t["press"]) - (kin + np.asarray(vir)) / V)) <= 1e-9 * scale
    engine.remove(sid)


@pytest.mark.parametrize("ensemble", [0, 1])
def test_thermo_largest_velocity_and_force(fmd, engine, ensemble):
    """The thermo record's maxV and maxF (max_vel_force_kernel,
    src/analysis/thermo.cu:26-75, printed by src/main.cu:645-656) are the
    largest |v| and |F| of one atom at that step, with that atom's global id:
    checked against the returned velocities and a fresh force evaluation at
    the returned positions."""
    d = relax_kg(fmd, engine, 12, 40, seed=79, ensemble=ensemble, T=(1.0, 1.0),
                 Tdamp=(1.0 if ensemble == 0 else 0.1), nsteps=100, thermo_freq=10)
    sid = engine.add(d)
    engine.step(sid, 30)
    rec, _ = engine.drain_thermo(sid)
    t = rec[-1]
    assert t["step"] == engine.steps_done(sid)
    p, v, _ = engine.state(sid)
    f, _, _ = engine.forces(sid)
    engine.remove(sid)
    sv = np.sqrt(np.sum(v.astype(np.float64) ** 2, axis=1))
    sf = np.sqrt(np.sum(f[:, :3].astype(np.float64) ** 2, axis=1))
    assert abs(t["max_v"] - sv.max()) <= 1e-6 * sv.max() and t["max_v_id"] == int(np.argmax(sv))
    assert abs(t["max_f"] - sf.max()) <= 1e-6 * sf.max() and t["max_f_id"] == int(np.argmax(sf))
    assert all(r["max_v"] > 0 and r["max_f"] > 0 for r in rec)
