"""Nose-Hoover chain, barostat and masses on the device against the pinned
oracle: the device chain on the known-answer stream of
tests/test_nose_hoover.cu:343-439, the extended-system state after NVT / NPT
runs, conservation of the extended energy on device runs
(tests/test_nose_hoover.cu:19-192), and mass scaling of the kicks and the
noise (tests/test_mass.cu:21-144)."""
import numpy as np
import pytest

from conftest import KG_AP, KG_BP, KG_LJ, lattice_chains, pos4_of

pytestmark = pytest.mark.gpu
F = np.float32


def test_device_chain_known_answer_stream(engine, orc, golden):
    """GPUChainMatchesHost: 10,000 KE draws (N = 100, T = 1.5, Tdamp = 0.5,
    dt = 0.002, M = 3), each applied as the post-force call with the raw KE
    and then as the next pre-force call with the KE it leaves (the carry),
    the chain state carried from draw to draw.  The device's FP64 chain
    equals the oracle's FP64 chain to rounding at every call, and stays
    within the reference's tolerances of its FP32 literal chain."""
    N, M, Td, T, dt = 100, 3, 0.5, 1.5, 0.002
    Q1, Qr = 3 * N * T * Td * Td, T * Td * Td
    ke = golden["nhke"].astype(np.float64)
    sc, xi_d, v_d = engine.nh_chain_run(M, Q1, Qr, np.zeros(10), np.zeros(10), ke, T, N, dt / 2, carry=True)
    xi = np.zeros(10)
    v = np.zeros(10)
    xi32 = np.zeros(10)
    v32 = np.zeros(10)
    ref = np.zeros(2 * len(ke))
    lit = np.zeros(4000)
    for t, k in enumerate(ke):
        s1, xi, v = orc.nh_chain(1, M, Q1, Qr, xi, v, k, T, N, dt / 2)
        s2, xi, v = orc.nh_chain(1, M, Q1, Qr, xi, v, k * s1 * s1, T, N, dt / 2)
        ref[2 * t], ref[2 * t + 1] = s1, s2
        if t < 2000:
            a1, xi32, v32 = orc.nh_chain(0, M, Q1, Qr, xi32, v32, k, T, N, dt / 2)
            a2, xi32, v32 = orc.nh_chain(0, M, Q1, Qr, xi32, v32, k * a1 * a1, T, N, dt / 2)
            lit[2 * t], lit[2 * t + 1] = a1, a2
    assert np.max(np.abs(sc - ref)) < 1e-14
    assert np.max(np.abs(v_d[:M] - v[:M])) < 1e-9 * max(1.0, np.max(np.abs(v[:M])))
    assert np.max(np.abs(xi_d[:M] - xi[:M])) < 1e-9 * max(1.0, np.max(np.abs(xi[:M])))
    assert np.all(v_d[M:] == 0) and np.all(xi_d[M:] == 0)
    # against the literal FP32 chain (nh_expf) of the reference, state carried
    assert np.max(np.abs(sc[:4000] - lit)) < 1e-6


def _nh_desc(fmd, p4, v, L, bonds, angles, ensemble, dt, Tdamp=0.1, P=1.0, Pdamp=1.0, nsteps=1000,
             masses=None, ntypes=1, lj=KG_LJ):
    return fmd.Desc(p4, L, lj, ntypes=ntypes, vel=v, bonds=bonds, bond_params=KG_BP, angles=angles,
                    angle_params=KG_AP, masses=masses, dt=dt, ensemble=ensemble, T=(1.0, 1.0), Tdamp=Tdamp,
                    P=(P, P), Pdamp=Pdamp, chain=3, nsteps=nsteps)


def _oracle_sys(p4, v, L, bonds, angles, ensemble, dt, Tdamp=0.1, P=1.0, Pdamp=1.0, nsteps=1000, masses=None,
                ntypes=1, lj=KG_LJ, seed=42):
    return {"pos4": p4.copy(), "vel3": v.copy(), "img3": np.zeros((len(p4), 3), np.int32),
            "masses": np.ones(16, F) if masses is None else np.resize(np.asarray(masses, F), 16), "lj": lj,
            "ntypes": ntypes, "bonds": bonds, "btype": np.zeros(len(bonds), np.int32), "bp": KG_BP,
            "ang4": angles, "ap": KG_AP, "L": float(L), "ensemble": ensemble, "T_start": 1.0, "T_stop": 1.0,
            "Tdamp": Tdamp, "P_start": P, "P_stop": P, "Pdamp": Pdamp, "chain": 3, "nsteps_total": nsteps,
            "dt": dt, "seed": seed}


@pytest.mark.parametrize("ensemble", [1, 2])
def test_extended_state_matches_oracle(fmd, engine, orc, ensemble):
    """40 NVT / NPT steps of a dense 343-bead KG system from a lattice start:
    the device's chain (xi, v_xi), eps, v_eps and carried KE equal the
    oracle's extended-system record of the same step."""
    p4, v, L, b, a = lattice_chains(7)
    sid = engine.add(_nh_desc(fmd, p4, v, L, b, a, ensemble, 0.002))
    th0 = engine.thermostat(sid)
    engine.step(sid, 40)
    th = engine.thermostat(sid)
    engine.remove(sid)
    *_, xr = orc.md_run(_oracle_sys(p4, v, L, b, a, ensemble, 0.002), 40, ext=True)
    # how far 40 steps carry a start error of the size of the device's FP32
    # force error: the oracle run again from positions moved by ~1e-6 sigma
    # (two FP32 roundings at this box size); the device may differ from the
    # oracle by a few times that spread, or by 2e-5 relative, whichever is larger
    pp = p4.copy()
    pp[:, :3] += F(1e-6) * np.random.default_rng(4).standard_normal((len(p4), 3)).astype(F)
    *_, xp = orc.md_run(_oracle_sys(pp, v, L, b, a, ensemble, 0.002), 40, ext=True)
    assert th0["Q1"] == xr[0, 4] and th0["Qr"] == xr[0, 5] and th0["W"] == xr[0, 6]
    assert th["step"] == 40 and th["M"] == 3
    x = xr[40]
    tol = lambda c, lo: max(2e-5 * max(lo, abs(x[c])), 4.0 * abs(xp[40, c] - x[c]))
    for k in range(3):
        assert abs(th["vxi"][k] - x[20 + k]) <= tol(20 + k, 1.0), (k, th["vxi"][k], x[20 + k], xp[40, 20 + k])
        assert abs(th["xi"][k] - x[10 + k]) <= tol(10 + k, 0.1), (k, th["xi"][k], x[10 + k], xp[40, 10 + k])
    if ensemble == 2:
        assert abs(th["eps"] - x[7]) <= 1e-6 + tol(7, 0.0)
        assert abs(th["veps"] - x[8]) <= tol(8, 1.0)
        assert np.max(np.abs(x[7])) > 0  # the barostat moved


@pytest.mark.parametrize("ensemble", [1, 2])
def test_extended_energy_conserved_on_device(fmd, engine, orc, ensemble):
    """The extended energy of the chain's (and barostat's) own dynamics stays
    within the drift the oracle shows for the same system and dt (pinned by
    its dt convergence in the oracle tests): 300 steps at dt = 0.002, read
    every 30 steps between runs; the LJ energy is taken cut and shifted."""
    p4, v, L, b, a = lattice_chains(7)
    n = len(p4)
    dt = 0.002
    sid = engine.add(_nh_desc(fmd, p4, v, L, b, a, ensemble, dt))
    rows = []
    for k in range(11):
        if k:
            engine.step(sid, 30)
        t = engine.thermo(sid)
        th = engine.thermostat(sid)
        pk, _, _ = engine.state(sid)
        Lp = engine.state_box(sid)
        x = np.zeros(orc.EXT)
        x[0], x[1] = t["ke"] * n, t["pe"] * n
        x[3] = th["V0"] * np.exp(3.0 * th["eps"])
        x[4], x[5], x[6], x[7], x[8], x[9] = th["Q1"], th["Qr"], th["W"], th["eps"], th["veps"], th["M"]
        x[10:20], x[20:30], x[30], x[31] = th["xi"], th["vxi"], 1.0, 1.0
        x[32] = orc.lj_cut_energy(pk, F(Lp), KG_LJ, bonds=b)
        rows.append(x)
    engine.remove(sid)
    xr = np.array(rows)
    # the integral term of the NPT form needs every step; read each 30 steps it
    # is taken by the trapezoid rule over the 30-step spacing
    H = orc.extended_energy(xr, n, 30 * dt)
    *_, xo = orc.md_run(_oracle_sys(p4, v, L, b, a, ensemble, dt), 300, ext=True)
    Ho = orc.extended_energy(xo, n, dt)
    drift_o = np.max(np.abs(Ho - Ho[0]))
    drift_d = np.max(np.abs(H - H[0]))
    assert drift_d < 1.5 * drift_o + 0.002 * xr[0, 0], (drift_d, drift_o)


@pytest.mark.parametrize("ensemble", [0, 1, 2])
def test_two_mass_trajectory_matches_oracle(fmd, engine, orc, ensemble):
    """tests/test_mass.cu: beads of masses 1 and 3 (alternating types with
    identical LJ), 40 steps of Langevin, NVT and NPT against the oracle
    driver with the same masses and noise: positions, velocities and the
    per-step temperature (mass-weighted KE)."""
    p4, v, L, b, a = lattice_chains(7, seed=4)
    types = np.arange(len(p4)) % 2
    p4 = pos4_of(p4[:, :3], types)
    masses = [1.0, 3.0]
    v = (v / np.sqrt(np.where(types == 0, 1.0, 3.0))[:, None]).astype(F)
    lj = np.ones(8, F)
    Tdamp = 1.0 if ensemble == 0 else 0.1
    d = _nh_desc(fmd, p4, v, L, b, a, ensemble, 0.002, Tdamp=Tdamp, masses=masses, ntypes=2, lj=lj)
    d.c.seed = 77
    d.c.thermo_freq = 1
    sid = engine.add(d)
    engine.step(sid, 40)
    pd, vd, _ = engine.state(sid)
    rec, _ = engine.drain_thermo(sid)
    engine.remove(sid)
    po, vo, _, Lo, th = orc.md_run(_oracle_sys(p4, v, L, b, a, ensemble, 0.002, Tdamp=Tdamp, masses=masses,
                                               ntypes=2, lj=lj, seed=77), 40)
    dx = pd[:, :3].astype(np.float64) - po[:, :3]
    dx -= Lo * np.rint(dx / Lo)
    assert np.max(np.abs(dx)) < 2e-4
    assert np.max(np.abs(vd - vo)) < 5e-3
    for k, r in enumerate(rec):
        assert abs(r["temp"] - th[k, 3]) <= 2e-4 * th[k, 3]


def test_half_kick_mass_scaling(fmd, engine, orc):
    """tests/test_mass.cu:21-63: two beads of masses 1 and 2 pushed apart by
    their LJ repulsion for one step (NVE limit): m1 dv1 = -m2 dv2, and both
    velocities equal the oracle driver's."""
    p4 = pos4_of([[5.0, 5.0, 5.0], [6.0, 5.0, 5.0]], [0, 1])
    d = fmd.Desc(p4, 12.0, np.ones(8, F), ntypes=2, vel=np.zeros((2, 3)), masses=[1.0, 2.0], Tdamp=1e10)
    sid = engine.add(d)
    engine.step(sid, 1)
    _, v, _ = engine.state(sid)
    engine.remove(sid)
    sysd = {"pos4": p4, "vel3": np.zeros((2, 3), F), "masses": np.array([1.0, 2.0] + [1.0] * 14, F),
            "lj": np.ones(8, F), "ntypes": 2, "L": 12.0, "Tdamp": 1e10, "ensemble": 0, "nsteps_total": 1000}
    _, vo, _, _, _ = orc.md_run(sysd, 1)
    assert v[0, 0] < 0 < v[1, 0]
    assert abs(v[0, 0] + 2.0 * v[1, 0]) < 1e-6 * abs(v[0, 0])
    assert np.max(np.abs(v - vo)) <= 1e-6 * abs(vo[0, 0])
