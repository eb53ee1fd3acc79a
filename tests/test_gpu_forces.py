"""Force, energy and virial parity of the device force kernel with the oracle:
the reference's own test inputs at the reference's tolerances against the
literal FP32 restatement of reference/cpu_reference.hpp, and against the FP64
oracle at 1e-5 relative (per-atom force error floored at the RMS force)."""
import numpy as np
import pytest

from conftest import KG_AP, KG_BP, KG_LJ, pos4_of, rel_err_floor, relax_kg

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative, against the FP64 oracle


def _f64_check(f_gpu, vir_gpu, pe_gpu, f8, vir8):
    fr = f8[:, :3]
    floor = np.sqrt(np.mean(np.sum(fr ** 2, axis=1)))
    assert rel_err_floor(f_gpu[:, :3], fr, floor) < TOL
    pe8 = f8[:, 3].sum()
    assert abs(pe_gpu - pe8) <= TOL * max(abs(pe8), 1e-3 * len(f8))
    scale = max(np.abs(vir8).max(), 1e-12)
    assert np.max(np.abs(vir_gpu - vir8)) <= TOL * scale
    np.testing.assert_allclose(f_gpu[:, 3], f8[:, 3], rtol=TOL, atol=TOL * np.abs(f8[:, 3]).max())


def _clamped_atoms(p4, L, rc):
    """Atoms in a pair whose LJ force/r exceeds the +-1000 clamp of
    src/force/lj.cu:48 (the reference CPU code does not clamp)."""
    x = p4[:, :3].astype(np.float64)
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.rint(d / L)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    with np.errstate(divide="ignore", over="ignore"):
        fr = 24.0 / r2 * (2.0 / r2 ** 6 - 1.0 / r2 ** 3)
    return np.any((r2 < rc * rc) & (np.abs(fr) > 1000.0), axis=1)


def test_lj_reference_input(fmd, engine, orc, golden):
    # tests/test_lj.cu:7-77
    p4 = pos4_of(golden["lj64_xyz"])
    L = np.float32(8.0)
    sid = engine.add(fmd.Desc(p4, L, KG_LJ, rc=2.5, skin=0.5))
    f, vir, pe = engine.forces(sid, fmd.TERM_LJ)
    fl, vl = orc.lj_literal(p4, KG_LJ, 1, 2.5, L)
    # the reference's own tolerance (tests/test_lj.cu:60) on every atom the
    # force/r clamp leaves alone; clamped atoms are checked against the clamped
    # FP64 oracle below
    free = ~_clamped_atoms(p4, float(L), 2.5)
    assert free.sum() >= 40
    assert rel_err_floor(f[free, :3], fl[free, :3], 1e-12) < 1e-3
    f8, v8 = orc.forces_f64(p4, L, KG_LJ, terms=1)
    _f64_check(f, vir, pe, f8, v8)
    engine.remove(sid)


def test_fene_reference_input(fmd, engine, orc):
    # tests/test_fene.cu:6-56: 3 atoms, k=30, R0=1.5, eps=sig=1
    p4 = pos4_of([[2.0, 5.0, 5.0], [3.0, 5.0, 5.0], [4.2, 5.0, 5.0]])
    bonds = np.array([[0, 1], [1, 2]], np.int32)
    L = np.float32(10.0)
    sid = engine.add(fmd.Desc(p4, L, [0.0, 0.0], bonds=bonds, bond_params=KG_BP))
    f, vir, pe = engine.forces(sid, fmd.TERM_BOND)
    fl, vl = orc.fene_literal(p4, bonds, [0, 0], KG_BP, L)
    assert rel_err_floor(f[:, :3], fl[:, :3], 1e-12) < 1e-3
    f8, v8 = orc.forces_f64(p4, L, [0.0, 0.0], bonds=bonds, bp=KG_BP, terms=2)
    _f64_check(f, vir, pe, f8, v8)
    # closed form at r = 1: -kR0^2/(R0^2-1) + 24 eps (2-1) = -30 along +x on atom 0
    assert abs(f8[0, 0] - (-30.0 * 2.25 / 1.25 + 24.0) * -1.0) < 1e-9
    engine.remove(sid)


def test_angle_reference_input(fmd, engine, orc):
    # tests/test_angle.cu:6-62: 90-degree triangle, k=100, theta0=120 deg
    p4 = pos4_of([[3.0, 5.0, 5.0], [5.0, 5.0, 5.0], [5.0, 7.0, 5.0]])
    ang = np.array([[0, 1, 2, 0]], np.int32)
    th0 = np.float32(np.pi * 120.0 / 180.0)
    ap = np.array([[100.0, th0]], np.float32)
    L = np.float32(10.0)
    sid = engine.add(fmd.Desc(p4, L, [0.0, 0.0], angles=ang, angle_params=ap))
    f, vir, pe = engine.forces(sid, fmd.TERM_ANGLE)
    fl, vl = orc.angle_literal(p4, ang, ap, L)
    assert rel_err_floor(f[:, :3], fl[:, :3], 1e-12) < 1e-3
    assert np.all(np.abs(f[:, :3].sum(axis=0)) < 1e-3)  # tests/test_angle.cu:53-58
    f8, v8 = orc.forces_f64(p4, L, [0.0, 0.0], ang4=ang, ap=ap, terms=4)
    _f64_check(f, vir, pe, f8, v8)
    # closed form: dtheta = -pi/6 (with theta0 rounded to FP32), |F_i| = 2k|dtheta|/|r_ij|
    dth = np.pi / 2 - float(th0)
    assert abs(f8[0, 1] - 2 * 100.0 * dth / 2.0) < 1e-9 and abs(f8[0, 0]) < 1e-12
    assert abs(pe - 100.0 * dth * dth) < 1e-4
    engine.remove(sid)


@pytest.mark.parametrize("nch,clen,seed", [(10, 100, 11), (40, 25, 12)])
def test_kg_melt_all_terms(fmd, engine, orc, nch, clen, seed):
    d = relax_kg(fmd, engine, nch, clen, seed=seed)
    sid = engine.add(d)
    engine.step(sid, 30)
    p4, _, _ = engine.state(sid)
    L = np.float32(d.c.L)
    f8, v8 = orc.forces_f64(p4, L, KG_LJ, bonds=d.bonds, btypes=d.bond_types, bp=KG_BP, ang4=d.angles,
                            ap=KG_AP, terms=7)
    for terms in (fmd.TERM_LJ, fmd.TERM_BOND, fmd.TERM_ANGLE, fmd.TERM_ALL):
        f, vir, pe = engine.forces(sid, terms)
        f8t, v8t = orc.forces_f64(p4, L, KG_LJ, bonds=d.bonds, btypes=d.bond_types, bp=KG_BP,
                                  ang4=d.angles, ap=KG_AP, terms=terms)
        _f64_check(f, vir, pe, f8t, v8t)
    # momentum: forces sum to zero (Newton's third law across all terms)
    f, vir, pe = engine.forces(sid, fmd.TERM_ALL)
    assert np.all(np.abs(f[:, :3].astype(np.float64).sum(axis=0)) < 1e-5 * np.abs(f8[:, :3]).sum() / len(f8) * 10)
    engine.remove(sid)


# three bead types with distinct LJ parameters per type pair (symmetric)
LJ3 = np.array([[1.0, 1.0], [0.75, 1.1], [1.2, 0.9],
                [0.75, 1.1], [0.5, 1.2], [0.9, 1.05],
                [1.2, 0.9], [0.9, 1.05], [1.5, 0.8]], np.float32).reshape(-1)


def test_lj_three_types_distinct_parameters(fmd, engine, orc):
    """Per-type-pair epsilon and sigma (three types, every pair different):
    a dense bonded system from a lattice start, forces/PE/virial against the
    FP64 oracle at 1e-5 and the literal FP32 restatement of
    reference/cpu_reference.hpp:95-153 (ti * ntypes + tj indexing of the
    parameter table), and 40 Langevin steps with per-type masses against the
    oracle driver."""
    from conftest import lattice_chains
    p4, v, L, b, a = lattice_chains(8, rho=0.7, seed=5)
    rng = np.random.default_rng(6)
    types = rng.integers(0, 3, len(p4))
    p4 = pos4_of(p4[:, :3], types)
    masses = np.array([1.0, 2.0, 0.5], np.float32)
    d = fmd.Desc(p4, L, LJ3, ntypes=3, vel=v, bonds=b, bond_params=KG_BP, angles=a, angle_params=KG_AP,
                 masses=masses, seed=11, T=(1.0, 1.0), Tdamp=1.0, nsteps=1000, thermo_freq=1)
    sid = engine.add(d)
    # the LJ term alone skips bonded pairs (they are in no Verlet row), so the
    # oracle is given the bonds for their exclusions
    f, vir, pe = engine.forces(sid, fmd.TERM_LJ)
    f8, v8 = orc.forces_f64(p4, L, LJ3, ntypes=3, bonds=b, terms=1)
    _f64_check(f, vir, pe, f8, v8)
    fl, _ = orc.lj_literal(p4, LJ3, 3, 2.5, L, excl=orc.exclusions(len(p4), b))
    floor = np.sqrt(np.mean(np.sum(f8[:, :3] ** 2, axis=1)))
    assert rel_err_floor(fl[:, :3], f8[:, :3], floor) < 1e-4
    f, vir, pe = engine.forces(sid)
    f8, v8 = orc.forces_f64(p4, L, LJ3, ntypes=3, bonds=b, bp=KG_BP, ang4=a, ap=KG_AP, terms=7)
    _f64_check(f, vir, pe, f8, v8)
    engine.step(sid, 40)
    pd, vd, _ = engine.state(sid)
    rec, _ = engine.drain_thermo(sid)
    engine.remove(sid)
    sysd = {"pos4": p4.copy(), "vel3": v.copy(), "masses": np.resize(masses, 16), "lj": LJ3, "ntypes": 3,
            "bonds": b, "btype": np.zeros(len(b), np.int32), "bp": KG_BP, "ang4": a, "ap": KG_AP, "L": float(L),
            "ensemble": 0, "seed": 11, "nsteps_total": 1000}
    po, vo, _, Lo, th = orc.md_run(sysd, 40)
    dx = pd[:, :3].astype(np.float64) - po[:, :3]
    dx -= Lo * np.rint(dx / Lo)
    assert np.max(np.abs(dx)) < 2e-4 and np.max(np.abs(vd - vo)) < 5e-3
    for k, r in enumerate(rec):
        assert abs(r["temp"] - th[k, 3]) <= 2e-4 * th[k, 3]
