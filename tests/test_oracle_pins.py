"""Pins the CPU oracle (CPU only).  Each check is independent of the formula
it checks: published known-answer vectors, closed forms, invariants, two
different algorithms for the same set, and finite differences of the energy."""
import numpy as np
import pytest

from conftest import KG_AP, KG_BP, KG_LJ, lattice_chains, pos4_of

F = np.float32


# --- Philox4x32-10 known-answer vectors (Random123 kat_vectors, Salmon et al. SC'11)
@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_kat(orc, ctr, key, out):
    assert orc.philox(ctr, key) == out


def test_normals_are_standard(orc):
    z = np.array([orc.normal3(7, g, s) for g in range(300) for s in range(1, 40)]).ravel()
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1.0) < 0.02
    assert abs(np.mean(z ** 4) - 3.0) < 0.15  # Gaussian kurtosis
    # independent streams per (gid, step)
    assert not np.array_equal(orc.normal3(7, 1, 1), orc.normal3(7, 2, 1))
    assert not np.array_equal(orc.normal3(7, 1, 1), orc.normal3(7, 1, 2))
    assert np.array_equal(orc.normal3(7, 3, 5), orc.normal3(7, 3, 5))


# --- PBC (tests/test_pbc.cu:22-93 inputs)
def test_min_image_hand_cases(orc):
    L = F(10.0)
    r1 = [[1.0, 2.0, 3.0], [9.5, 0.5, 5.0], [0.5, 0.5, 0.5], [5.0, 5.0, 5.0], [0.1, 9.9, 5.0]]
    r2 = [[2.0, 3.0, 4.0], [0.5, 9.5, 5.0], [9.5, 9.5, 9.5], [5.0, 5.0, 5.0], [9.9, 0.1, 5.0]]
    exact = [[-1, -1, -1], [-1, 1, 0], [1, 1, 1], [0, 0, 0], [0.2, -0.2, 0.0]]
    for a, b, e in zip(r1, r2, exact):
        for d in range(3):
            got = orc.min_image(F(F(a[d]) - F(b[d])), L)
            assert abs(got - e[d]) <= 4e-7 * 10 and -5.0 <= got <= 5.0


def test_wrap_lands_in_box(orc):
    L = F(10.0)
    for x in [-0.5, 10.5, 5.0, 20.3, -3.7, 0.0, 10.0, -0.01, 1e-8, -1e-8]:
        w = orc.wrap(F(x), L)
        assert 0.0 <= w <= L  # == L only for tiny negatives rounding up
        assert abs(((x - w) / 10.0) - round((x - w) / 10.0)) < 1e-6


@pytest.mark.parametrize("L", [8.0, 33.0, 105.567])
def test_fixed_point_displacement(orc, L):
    """The force-path displacement is the minimum-image difference of positions
    quantised to L/2^32, rounded twice in FP32: within one grid step plus 2 ulp
    of the exact minimum-image difference, and antisymmetric."""
    rng = np.random.default_rng(int(L))
    L = F(L)
    a = (rng.random(3000) * L).astype(F)
    b = (rng.random(3000) * L).astype(F)
    a[:4] = [0.0, np.nextafter(L, 0, dtype=F), F(1e-7), L / 2]
    b[:4] = [np.nextafter(L, 0, dtype=F), 0.0, L - F(1e-7), 0.0]
    q = float(L) / 2 ** 32
    for x, y in zip(a, b):
        got = orc.lib().orc_min_image_fixed(x, y, L)
        d = float(x) - float(y)
        d -= float(L) * np.rint(d / float(L))
        if abs(abs(d) - float(L) / 2) < 1e-4:
            continue  # the two images are equally near
        assert abs(got - d) <= 2 * q + 2 * abs(d) * 2 ** -24
        assert orc.lib().orc_min_image_fixed(y, x, L) == -got or abs(got) < 3 * q


def test_min_image_single_rounding_form_is_exact():
    """The list build evaluates dr - L * rint(dr / L) as fma(-L, m, dr): for
    wrapped positions m is -1, 0 or 1, L * m is exact, so one rounding of the
    exact dr - L m equals the reference's two-step FP32 result.  Checked with
    FP64 (where dr - L m is exact) rounded once to FP32."""
    rng = np.random.default_rng(3)
    for L in (F(6.0), F(13.3), F(33.6), F(105.567)):
        a = (rng.random(200000) * L).astype(F)
        b = (rng.random(200000) * L).astype(F)
        a[:4] = [0.0, np.nextafter(L, 0, dtype=F), L, F(1e-7)]
        b[:4] = [np.nextafter(L, 0, dtype=F), 0.0, 0.0, L]
        dr = (a - b).astype(F)
        m = np.rint(dr * (F(1) / L)).astype(F)
        assert set(np.unique(m)) <= {F(-1), F(0), F(1)}
        two_step = (dr - (L * m).astype(F)).astype(F)
        fused = (dr.astype(np.float64) - np.float64(L) * m.astype(np.float64)).astype(F)
        assert np.array_equal(two_step.view(np.uint32), fused.view(np.uint32))


@pytest.mark.parametrize("L", [6.0, 31.7, 105.6])
def test_lj_grid_cutoff_matches_exact_distance(orc, L):
    """The device's grid-unit LJ test equals the exact minimum-image test
    r^2 < rc^2 (and r^2 > 1e-10) outside a thin band around rc^2: a wrong
    scale of the grid (u instead of u^2, a missing 1/L) or a dropped axis
    flips many pairs far from the band."""
    rng = np.random.default_rng(7)
    L = F(L)
    rc2 = F(2.5) * F(2.5)
    n = 20000
    a = (rng.random((n, 3)) * L).astype(F)
    # offsets spread around the cutoff, and a few coincident and near pairs
    r = rng.uniform(0.0, 3.2, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    b = (a + r[:, None] * u).astype(F)
    b[:5] = a[:5]
    b[5:10] = a[5:10] + F(2e-5)
    b = np.mod(b, L).astype(F)
    got = orc.lj_inside_grid(a, b, float(rc2), float(L))
    d = a.astype(np.float64) - b.astype(np.float64)
    d -= float(L) * np.rint(d / float(L))
    r2 = (d * d).sum(1)
    want = (r2 < float(rc2)) & (r2 > 1e-10)
    band = np.abs(r2 - float(rc2)) < 1e-5 * float(rc2)
    assert not got[:5].any()              # coincident points (row padding) are never a pair
    assert got[5:10].all()                # 3.5e-5 apart (r^2 = 1.2e-9): inside
    assert np.array_equal(got[~band], want[~band])
    assert band.sum() < 20 and want.sum() > n // 2


# --- Morton keys: decode the interleave bit by bit and compare with the quantised coordinates
def _deinterleave(k):
    out = [0, 0, 0]
    for bit in range(10):
        for a in range(3):
            out[a] |= ((k >> (3 * bit + a)) & 1) << bit
    return out


def test_morton_keys_decode(orc, golden):
    L = F(10.0)
    for x, y, z in golden["morton64_xyz"]:
        k = orc.morton3d(x, y, z, L)
        q = [min(max(int(np.floor(float(F(F(c) * (F(1) / L)) * F(1024)))), 0), 1023) for c in (x, y, z)]
        assert _deinterleave(k) == q
    assert orc.morton3d(F(9.9999), F(9.9999), F(9.9999), L) == (1 << 30) - 1


def test_morton_perm_is_stable_sort(orc, golden):
    p4 = pos4_of(golden["morton64_xyz"])
    keys = [orc.morton3d(*x, F(10.0)) for x in golden["morton64_xyz"]]
    assert list(orc.morton_perm(p4, F(10.0))) == sorted(range(64), key=lambda i: (keys[i], i))


# --- cells: every atom lies in the bounds of its cell
@pytest.mark.parametrize("L,rcs", [(32.802, 2.8), (8.0, 3.0), (13.3, 2.8), (105.567, 2.8)])
def test_cell_ids_contain_atoms(orc, L, rcs):
    rng = np.random.default_rng(1)
    p4 = pos4_of((rng.random((3000, 3)) * L).astype(F))
    cid, nx = orc.cell_ids(p4, F(L), F(rcs))
    assert nx == max(1, int(F(L) / F(rcs)))
    cs = L / nx
    c = np.stack([cid % nx, (cid // nx) % nx, cid // (nx * nx)], 1)
    x = p4[:, :3].astype(np.float64)
    assert np.all(x >= c * cs - 1e-4) and np.all(x < (c + 1) * cs + 1e-4)


# --- pair sets: brute force and cell enumeration agree; cutoff is strict
@pytest.mark.parametrize("n,L,rcs,bonded", [(800, 9.8, 2.8, False), (600, 7.0, 3.0, True), (300, 5.5, 2.8, True),
                                            (900, 15.0, 2.8, True)])
def test_pair_set_brute_equals_cells(orc, n, L, rcs, bonded):
    rng = np.random.default_rng(n)
    p4 = pos4_of((rng.random((n, 3)) * L).astype(F))
    excl = None
    if bonded:
        bonds = np.stack([np.arange(0, n - 1, 2), np.arange(1, n, 2)], 1)
        excl = orc.exclusions(n, bonds)
    a = orc.pair_set(p4, F(L), F(rcs), excl, method=0)
    b = orc.pair_set(p4, F(L), F(rcs), excl, method=1)
    assert np.array_equal(a, b) and len(a) > 0
    # symmetric, no self pairs, excluded pairs absent
    s = set(map(tuple, a.tolist()))
    assert all((j, i) in s for i, j in s) and all(i != j for i, j in s)
    if bonded:
        assert (0, 1) not in s and (1, 0) not in s


def test_pair_cutoff_strict(orc):
    # two atoms exactly rcs apart along x: not neighbors (strict <, verlet_list.cu:268)
    p4 = pos4_of([[1.0, 5.0, 5.0], [4.0, 5.0, 5.0]])
    assert len(orc.pair_set(p4, F(20.0), F(3.0))) == 0
    p4 = pos4_of([[1.0, 5.0, 5.0], [3.99, 5.0, 5.0]])
    assert len(orc.pair_set(p4, F(20.0), F(3.0))) == 2


def test_exclusions_csr(orc):
    off, lst = orc.exclusions(5, [[0, 1], [1, 2], [2, 1], [3, 4]])
    assert list(off) == [0, 1, 3, 4, 5, 6] and list(lst) == [1, 0, 2, 1, 4, 3]


# --- forces: closed forms and finite differences of the FP64 energy
def _energy(orc, p4, L, terms, **kw):
    f8, _ = orc.forces_f64(p4, F(L), KG_LJ, terms=terms, clamp_lj=0, clamp_fene=0, **kw)
    return f8[:, 3].sum()


def test_lj_closed_forms(orc):
    rmin = 2 ** (1 / 6)
    p4 = pos4_of([[5.0, 5.0, 5.0], [5.0 + rmin, 5.0, 5.0]])
    f8, vir = orc.forces_f64(p4, F(20.0), KG_LJ, terms=1)
    # force vanishes at the minimum, up to U''(rmin) = 57.1 times the FP32 rounding of r
    dr = abs(float(p4[1, 0]) - float(p4[0, 0]) - rmin)
    assert np.all(np.abs(f8[:, :3]) <= 57.2 * dr + 1e-9)
    assert abs(f8[:, 3].sum() + 1.0) < 1e-6           # pair energy -eps, half to each
    p4 = pos4_of([[5.0, 5.0, 5.0], [6.0, 5.0, 5.0]])
    f8, vir = orc.forces_f64(p4, F(20.0), KG_LJ, terms=1)
    assert abs(f8[0, 0] + 24.0) < 1e-9 and abs(f8[1, 0] - 24.0) < 1e-9
    assert abs(vir[0] - 24.0) < 1e-9                  # W_xx = f.r for the pair (half from each atom)
    assert abs(f8[:, 3].sum()) < 1e-12


def test_fene_wca_closed_form(orc):
    p4 = pos4_of([[5.0, 5.0, 5.0], [6.0, 5.0, 5.0]])
    f8, vir = orc.forces_f64(p4, F(20.0), KG_LJ, bonds=[[0, 1]], bp=KG_BP, terms=2)
    ff = -30.0 * 2.25 / (2.25 - 1.0) + 24.0
    assert abs(f8[0, 0] - ff * -1.0) < 1e-9
    pe = -0.5 * 30.0 * 2.25 * np.log(1 - 1 / 2.25) + 1.0
    assert abs(f8[:, 3].sum() - pe) < 1e-9


def test_angle_straight_chain_is_force_free(orc):
    p4 = pos4_of([[4.0, 5.0, 5.0], [5.0, 5.0, 5.0], [6.0, 5.0, 5.0]])
    f8, _ = orc.forces_f64(p4, F(20.0), KG_LJ, ang4=[[0, 1, 2, 0]], ap=KG_AP, terms=4)
    assert np.all(np.abs(f8[:, :3]) < 1e-5) and abs(f8[:, 3].sum()) < 1e-12


def _saw_chain(n, L, seed, bond=0.97, dmin=0.95):
    """A self-avoiding random-walk chain (all non-bonded distances >= dmin)."""
    rng = np.random.default_rng(seed)
    x = [np.full(3, L / 2)]
    while len(x) < n:
        u = rng.normal(size=3)
        c = x[-1] + bond * u / np.linalg.norm(u)
        if len(x) >= 2:
            d = np.array(x[:-1]) - c
            d -= L * np.rint(d / L)
            if np.min(np.linalg.norm(d, axis=1)) < dmin:
                continue
        x.append(c)
    xyz = np.mod(np.array(x), L).astype(F)
    bonds = np.array([[i, i + 1] for i in range(n - 1)], np.int32)
    ang = np.array([[i, i + 1, i + 2, 0] for i in range(n - 2)], np.int32)
    return xyz, dict(bonds=bonds, bp=KG_BP, ang4=ang, ap=KG_AP)


@pytest.mark.parametrize("terms", [1, 2, 4, 7])
def test_forces_are_minus_energy_gradient(orc, terms):
    L = 7.0
    xyz, kw = _saw_chain(60, L, seed=terms)
    p4 = pos4_of(xyz)
    f8, _ = orc.forces_f64(p4, F(L), KG_LJ, terms=terms, clamp_lj=0, clamp_fene=0, **kw)
    rng = np.random.default_rng(100 + terms)
    h = 1e-3
    for i in rng.choice(len(xyz), 10, replace=False):
        for d in range(3):
            pp, pm = p4.copy(), p4.copy()
            pp[i, d] = F(p4[i, d] + h)
            pm[i, d] = F(p4[i, d] - h)
            dx = float(pp[i, d]) - float(pm[i, d])
            fd = -(_energy(orc, pp, L, terms, **kw) - _energy(orc, pm, L, terms, **kw)) / dx
            assert abs(fd - f8[i, d]) <= 2e-3 * max(1.0, abs(f8[i, d]))


@pytest.mark.parametrize("terms", [1, 2, 4, 7])
def test_virial_trace_is_volume_derivative(orc, terms):
    """W_xx+W_yy+W_zz = -dU/dlambda at lambda=1 for r -> lambda r, L -> lambda L."""
    L = 7.0
    xyz, kw = _saw_chain(60, L, seed=10 + terms)
    _, vir = orc.forces_f64(pos4_of(xyz), F(L), KG_LJ, terms=terms, clamp_lj=0, clamp_fene=0, **kw)
    if terms == 4:
        # a bending energy depends on angles only, so it is scale-invariant: zero trace
        assert abs(vir[0] + vir[3] + vir[5]) < 1e-9 * max(1.0, np.abs(vir).max())
        return
    # small h: the unshifted LJ energy jumps by 0.0163 when a pair crosses rc,
    # and no pair of these chains crosses within +-1e-5 scaling
    h = 1e-5
    ep = _energy(orc, pos4_of(xyz.astype(np.float64) * (1 + h)), L * (1 + h), terms, **kw)
    em = _energy(orc, pos4_of(xyz.astype(np.float64) * (1 - h)), L * (1 - h), terms, **kw)
    dU = (ep - em) / (2 * h)
    tr = vir[0] + vir[3] + vir[5]
    assert abs(tr + dU) <= 1e-2 * max(1.0, abs(tr))


def test_literal_lj_matches_f64(orc, golden):
    # the literal FP32 restatement and the FP64 yardstick agree on the reference
    # LJ input away from clamped pairs (no pair of this input is within r < 0.79)
    p4 = pos4_of(golden["lj64_xyz"])
    fl, vl = orc.lj_literal(p4, KG_LJ, 1, 2.5, F(8.0))
    f8, v8 = orc.forces_f64(p4, F(8.0), KG_LJ, terms=1, clamp_lj=0)
    err = np.linalg.norm(fl[:, :3] - f8[:, :3], axis=1) / np.maximum(np.linalg.norm(f8[:, :3], axis=1), 1.0)
    assert err.max() < 1e-3


def test_literal_lj_three_types_matches_f64(orc):
    """Three types with distinct (eps, sigma) per type pair: the literal FP32
    brute force and the FP64 cell-list path index the table as
    lj[ti * ntypes + tj] and agree; swapping two types' parameters changes
    the forces (the table is read per pair)."""
    lj3 = np.array([[1.0, 1.0], [0.75, 1.1], [1.2, 0.9], [0.75, 1.1], [0.5, 1.2], [0.9, 1.05],
                    [1.2, 0.9], [0.9, 1.05], [1.5, 0.8]], F).reshape(-1)
    p4, _, L, _, _ = lattice_chains(6, rho=0.7, seed=2)
    types = np.random.default_rng(3).integers(0, 3, len(p4))
    p4 = pos4_of(p4[:, :3], types)
    fl, vl = orc.lj_literal(p4, lj3, 3, 2.5, L)
    f8, v8 = orc.forces_f64(p4, L, lj3, ntypes=3, terms=1)
    assert np.allclose(fl, f8, rtol=1e-4, atol=1e-4 * np.abs(f8).max())
    assert np.allclose(vl, v8, rtol=1e-4, atol=1e-4 * np.abs(v8).max())
    sw = lj3.reshape(3, 3, 2)[[1, 0, 2]][:, [1, 0, 2]].reshape(-1)
    fs, _ = orc.forces_f64(p4, L, sw, ntypes=3, terms=1)
    assert np.max(np.abs(fs - f8)) > 1e-2 * np.abs(f8).max()


def test_literal_bonded_match_f64(orc):
    p4 = pos4_of([[2.0, 5.0, 5.0], [3.0, 5.0, 5.0], [4.2, 5.0, 5.0]])
    fl, vl = orc.fene_literal(p4, [[0, 1], [1, 2]], [0, 0], KG_BP, F(10.0))
    f8, v8 = orc.forces_f64(p4, F(10.0), KG_LJ, bonds=[[0, 1], [1, 2]], bp=KG_BP, terms=2)
    assert np.allclose(fl, f8, rtol=1e-5, atol=1e-4) and np.allclose(vl, v8, rtol=1e-5, atol=1e-4)
    ap = np.array([[100.0, F(np.pi * 120 / 180)]], F)
    p4 = pos4_of([[3.0, 5.0, 5.0], [5.0, 5.0, 5.0], [5.0, 7.0, 5.0]])
    fl, vl = orc.angle_literal(p4, [[0, 1, 2, 0]], ap, F(10.0))
    f8, v8 = orc.forces_f64(p4, F(10.0), KG_LJ, ang4=[[0, 1, 2, 0]], ap=ap, terms=4)
    assert np.allclose(fl, f8, rtol=1e-5, atol=1e-4)


# --- Nose-Hoover chain
def test_nh_chain_f32_literal_matches_f64(orc, golden):
    """tests/test_nose_hoover.cu:343-439 style: 10000 KE draws, FP32 literal
    chain (nh_expf Taylor) against the FP64 chain, state carried across trials."""
    N, M, Td, T, dt = 100, 3, 0.5, 1.5, 0.002
    Nf = 3 * N
    Q1, Qr = Nf * T * Td * Td, T * Td * Td
    xi32 = np.zeros(10); v32 = np.zeros(10)
    xi64 = np.zeros(10); v64 = np.zeros(10)
    for ke in golden["nhke"][:2000]:
        s32, xi32, v32 = orc.nh_chain(0, M, Q1, Qr, xi32, v32, float(ke), T, N, dt / 2)
        s64, xi64, v64 = orc.nh_chain(1, M, Q1, Qr, xi64, v64, float(ke), T, N, dt / 2)
        assert abs(s32 - s64) < 1e-6
    assert np.allclose(v32[:M], v64[:M], atol=1e-4)


def test_nh_chain_single_thermostat_closed_form(orc):
    # M = 1: v += dt4*G twice around exp(-dt2 v) with G recomputed from the scaled KE
    Nf, T, KE, Q, h = 30.0, 1.0, 20.0, 3.0, 0.01
    s, xi, v = orc.nh_chain(1, 1, Q, Q, np.zeros(10), np.zeros(10), KE, T, 10, h)
    g0 = (2 * KE - Nf * T) / Q
    v1 = g0 * h / 2
    s_ref = np.exp(-h * v1)
    g1 = (2 * KE * s_ref ** 2 - Nf * T) / Q
    assert abs(s - s_ref) < 1e-15 and abs(v[0] - (v1 + g1 * h / 2)) < 1e-15 and abs(xi[0] - h * v1) < 1e-15


def test_oracle_md_nve_conserves_energy(orc, golden):
    # the oracle driver's NVE limit on tests/test_integrator.cu's lattice
    sysd = {"pos4": pos4_of(golden["nve64_xyz"]), "vel3": golden["nve64_vel"].copy(),
            "img3": np.zeros((64, 3), np.int32), "masses": np.ones(16, F), "lj": KG_LJ, "L": 8.0,
            "skin": 0.5, "Tdamp": 1e10, "ensemble": 0}
    _, _, _, _, th = orc.md_run(sysd, 300)
    e = (th[:, 1] + th[:, 2]) * 64
    assert np.max(np.abs(e - e[0])) < 0.01 * abs(e[0])


# --- Nose-Hoover conserved quantities
def _nh_sys(p4, v, L, bonds, angles, ensemble, dt, nsteps, Tdamp=0.1, P=1.0, Pdamp=1.0, masses=None):
    return {"pos4": p4.copy(), "vel3": v.copy(), "img3": np.zeros((len(p4), 3), np.int32),
            "masses": np.ones(16, F) if masses is None else masses, "lj": KG_LJ,
            "bonds": bonds, "btype": np.zeros(len(bonds), np.int32), "bp": KG_BP, "ang4": angles, "ap": KG_AP,
            "L": float(L), "ensemble": ensemble, "T_start": 1.0, "T_stop": 1.0, "Tdamp": Tdamp,
            "P_start": P, "P_stop": P, "Pdamp": Pdamp, "chain": 3, "nsteps_total": nsteps, "dt": dt}


def test_nh_reference_conserved_quantity(orc, golden):
    """tests/test_nose_hoover.cu:43-192 (NVTLimitEnergyConservation): the
    64-atom lattice of tests/test_integrator.cu, L=8, rc 2.5, skin 0.5,
    dt 0.001, M=3, Tdamp=1000, 500 steps; the reference's conserved quantity
    KE + PE + 1/2 Q1 v0^2 + Nf kT xi0 + sum (1/2 Qr vk^2 + kT xik)
    (tests/test_nose_hoover.cu:19-40) drifts by less than 1%."""
    sysd = {"pos4": pos4_of(golden["nve64_xyz"]), "vel3": golden["nve64_vel"].copy(),
            "img3": np.zeros((64, 3), np.int32), "masses": np.ones(16, F), "lj": KG_LJ, "L": 8.0,
            "skin": 0.5, "Tdamp": 1000.0, "ensemble": 1, "chain": 3, "nsteps_total": 500}
    *_, xr = orc.md_run(sysd, 500, ext=True)
    H = orc.extended_energy(xr, 64, 0.001, form=0, shifted=False)
    assert abs(H[-1] - H[0]) / abs(H[0]) < 0.01
    # the chain masses of NoseHooverState::init (src/integrate/nose_hoover.cu:336-339)
    assert xr[0, 4] == 3 * 64 * 1.0 * 1000.0 ** 2 and xr[0, 5] == 1000.0 ** 2


def test_lj_cut_energy_closed_form(orc):
    """The cutoff shift used by the conservation checks: one LJ energy at rc
    per non-excluded pair inside rc, nothing for a pair outside or a bonded
    (excluded) pair."""
    e_rc = 4.0 * (2.5 ** -12 - 2.5 ** -6)
    p4 = pos4_of([[1.0, 1.0, 1.0], [3.0, 1.0, 1.0], [1.0, 3.6, 1.0], [5.0, 5.0, 5.0]])
    assert abs(orc.lj_cut_energy(p4, F(10.0), KG_LJ) - e_rc) < 1e-15
    assert orc.lj_cut_energy(p4, F(10.0), KG_LJ, bonds=[[0, 1]]) == 0.0
    p4 = pos4_of([[0.5, 1.0, 1.0], [9.0, 1.0, 1.0]])  # 1.5 apart through the boundary
    assert abs(orc.lj_cut_energy(p4, F(10.0), KG_LJ) - e_rc) < 1e-15


def _drifts(orc, ensemble, t_end=0.5, form=1, shifted=True, **kw):
    p4, v, L, b, a = lattice_chains(7)
    out = []
    for dt in (0.002, 0.001):
        ns = int(round(t_end / dt))
        *_, xr = orc.md_run(_nh_sys(p4, v, L, b, a, ensemble, dt, ns, **kw), ns, ext=True)
        H = orc.extended_energy(xr, len(p4), dt, form=form, shifted=shifted)
        out.append((np.max(np.abs(H - H[0])), xr))
    return out


def test_nvt_extended_energy_converges(orc):
    """Moderate damping (Tdamp = 0.1, M = 3) on a dense 343-bead KG system
    with every force term: the extended energy of the chain's own dynamics
    (element k >= 1 weighted 2^k, oracle.hpp nh_chain_energy_dyn) is a
    constant of the continuous dynamics, so its drift over a fixed time
    falls with dt, at the first order of this chain splitting (the backward
    sweep leaves elements k >= 1 without their force term).  A plausible
    slip fails this: weighting the chain as the reference's test formula does
    leaves a drift of about 150 that does not shrink with dt, and so does an
    energy that ignores the jump of the unshifted LJ at the cutoff."""
    (d1, xr), (d2, _) = _drifts(orc, 1)
    K0 = xr[0, 0]
    assert 0.4 < d2 / d1 < 0.62, (d1, d2)
    assert d2 < 0.01 * K0
    (w1, _), (w2, _) = _drifts(orc, 1, form=0)
    assert w2 > 10 * d2 and w2 / w1 > 0.9
    (u1, _), (u2, _) = _drifts(orc, 1, shifted=False)
    assert u2 / u1 > 0.9
    # Q1 = Nf kT Tdamp^2, Qr = kT Tdamp^2 (src/integrate/nose_hoover.cu:336-339)
    tau2 = float(F(0.1)) ** 2  # Tdamp is an FP32 run.conf value
    assert abs(xr[0, 4] - 3 * 343 * tau2) < 1e-12 and abs(xr[0, 5] - tau2) < 1e-15


def test_npt_extended_energy_converges(orc):
    """NPT-NH at P = 1 (Pdamp = 1, Tdamp = 0.1, M = 3) from the same state:
    K + U(shifted) + P V + v_eps^2/(2W) + chain + the integral of the missing
    MTK term (6K/Nf) v_eps/W is a constant of the reference's barostat
    dynamics; its drift falls with dt at first order.  The barostat moves:
    eps swings by several percent of its range over the window."""
    (d1, xr), (d2, _) = _drifts(orc, 2, t_end=0.6)
    assert 0.4 < d2 / d1 < 0.62, (d1, d2)
    assert d2 < 0.03 * xr[0, 0]
    assert np.max(np.abs(xr[:, 7])) > 0.02
    # W = (Nf + 3) kT Pdamp^2 (src/integrate/nose_hoover.cu:341-343)
    assert abs(xr[0, 6] - (3 * 343 + 3)) < 1e-9


def test_two_mass_langevin_equipartition(orc):
    """tests/test_mass.cu restated through the driver: an ideal gas (LJ
    epsilon 0) of beads with masses 1 and 4 under Langevin at T = 1.5 reaches
    <m v^2> = kT per component for both species (the kick's 1/m and the
    noise's sqrt(kT/m), src/integrate/langevin.cu:48-66); KE is mass
    weighted (the thermo tensor sum m v v)."""
    n = 2000
    rng = np.random.default_rng(3)
    xyz = rng.uniform(0, 60.0, (n, 3))
    types = np.arange(n) % 2
    sysd = {"pos4": pos4_of(xyz, types), "vel3": np.zeros((n, 3), F), "masses": np.array([1.0, 4.0] + [1.0] * 14, F),
            "lj": np.zeros(8, F), "ntypes": 2, "L": 60.0, "Tdamp": 0.05, "T_start": 1.5, "T_stop": 1.5,
            "ensemble": 0, "nsteps_total": 600, "dt": 0.002}
    _, v, _, _, th = orc.md_run(sysd, 600)
    for t, m in ((0, 1.0), (1, 4.0)):
        mv2 = m * np.mean(v[types == t].astype(np.float64) ** 2)
        assert abs(mv2 - 1.5) < 0.06, (t, mv2)
    ke = 0.5 * np.sum(np.where(types == 0, 1.0, 4.0)[:, None] * v.astype(np.float64) ** 2) / n
    assert abs(th[-1, 1] - ke) < 1e-9 * ke


def test_two_mass_momentum_exchange(orc):
    """tests/test_mass.cu:21-63: the half kicks are dt/2 F/m.  Two beads of
    masses 1 and 2 pushed apart by their LJ repulsion for one step of the
    driver (NVE limit): both kicks come from the same equal and opposite
    forces, so m1 dv1 = -m2 dv2 and dv1 = -2 dv2 (to FP32 rounding)."""
    p4 = pos4_of([[5.0, 5.0, 5.0], [6.0, 5.0, 5.0]], [0, 1])
    sysd = {"pos4": p4, "vel3": np.zeros((2, 3), F), "masses": np.array([1.0, 2.0] + [1.0] * 14, F),
            "lj": np.ones(8, F), "ntypes": 2, "L": 12.0, "Tdamp": 1e10, "ensemble": 0, "nsteps_total": 1}
    _, v, _, _, _ = orc.md_run(sysd, 1)
    assert v[0, 0] < 0 < v[1, 0]
    assert abs(v[0, 0] + 2.0 * v[1, 0]) < 1e-6 * abs(v[0, 0])
    assert np.all(v[:, 1:] == 0)


def test_literal_cpu_path_follows_the_driver(orc):
    """The single-core literal path (brute-force FP32 forces of
    reference/cpu_reference.hpp and a serial BAOAB step) is timed as the
    reference's CPU baseline; it runs the same trajectory as the driver (whose
    forces are FP64) up to FP32 rounding over 10 Langevin steps of a dense
    343-bead KG system."""
    p4, v, L, b, a = lattice_chains(7)
    n = len(p4)
    pl, vl, _ = orc.literal_md_langevin(p4, v, np.zeros((n, 3), np.int32), np.ones(16, F), float(L), 10, KG_LJ, b,
                                        np.zeros(len(b), np.int32), KG_BP, a, KG_AP, seed=9)
    sysd = {"pos4": p4.copy(), "vel3": v.copy(), "masses": np.ones(16, F), "lj": KG_LJ, "bonds": b,
            "btype": np.zeros(len(b), np.int32), "bp": KG_BP, "ang4": a, "ap": KG_AP, "L": float(L),
            "ensemble": 0, "seed": 9, "nsteps_total": 10}
    po, vo, _, _, _ = orc.md_run(sysd, 10)
    d = pl[:, :3].astype(np.float64) - po[:, :3]
    d -= float(L) * np.rint(d / float(L))
    assert np.max(np.abs(d)) < 1e-5 and np.max(np.abs(vl - vo)) < 1e-4
    assert np.max(np.abs(vl - v)) > 0.01  # it moved


# --- radius of gyration (tests/test_rg.cu:5-77)
def test_rg_reference_cases(orc):
    p = np.zeros((8, 4), F)
    p[:, 0] = [1, 3, 5, 7, 2, 4, 6, 8]
    assert abs(orc.rg(p, np.zeros((8, 3), np.int32), [0, 0, 0, 0, 1, 1, 1, 1], 10.0) - 2.23607) < 1e-5
    p = np.zeros((3, 4), F)
    p[:, 0] = [9, 1, 3]
    im = np.zeros((3, 3), np.int32)
    im[1:, 0] = 1
    assert abs(orc.rg(p, im, [0, 0, 0], 10.0) - 1.63299) < 1e-5


def test_rg_equals_pairwise_form(orc):
    """Rg^2 = sum_ij |r_i - r_j|^2 / (2 N^2) per chain, an independent form;
    and Rg is unchanged when a bead is shifted by a box vector with its image
    count adjusted."""
    rng = np.random.default_rng(4)
    L = 12.0
    n, nch = 60, 3
    x = np.cumsum(rng.normal(size=(n, 3)) * 0.5, axis=0) + 3.0
    mol = np.repeat(np.arange(nch), n // nch)
    img = np.floor(x / L).astype(np.int32)
    p = np.zeros((n, 4), F)
    p[:, :3] = (x - img * L).astype(F)
    unw = p[:, :3].astype(np.float64) + img * L
    want = []
    for c in range(nch):
        y = unw[mol == c]
        d = y[:, None, :] - y[None, :, :]
        want.append(np.sqrt((d ** 2).sum() / (2 * len(y) ** 2)))
    assert abs(orc.rg(p, img, mol, L) - np.mean(want)) < 1e-9 * np.mean(want) + 1e-6
    img2 = img.copy()
    img2[5, 1] += 1
    p2 = p.copy()
    p2[5, 1] = F(p[5, 1] - L)
    assert abs(orc.rg(p2, img2, mol, L) - orc.rg(p, img, mol, L)) < 1e-5


# ---------------------------------------------------------------------------
# Tabulated pairs (src/force/table.cu:6-182)
# ---------------------------------------------------------------------------
def _nodes():
    r = 0.5 + 0.125 * np.arange(17)          # exact binary nodes on [0.5, 2.5]
    return r, np.stack([r, 3.0 - r, r * r], axis=1).astype(F)  # rows (r, e, f)


def test_table_interpolant_closed_form(orc):
    """The linear interpolant of f(r) = r^2 on the segment [r_k, r_k+1] holding
    r is r^2 + (r - r_k)(r_k+1 - r); e(r) = 3 - r is reproduced exactly; r is
    clamped to the table's range.  Segments are located by searchsorted."""
    nodes, rows = _nodes()
    rng = np.random.default_rng(5)
    r = np.concatenate([rng.uniform(0.2, 2.8, 4000), nodes, [0.0, 0.49, 2.51, 9.0]])
    f, e = orc.table_eval(0.5, 2.5, rows, r)
    rc = np.clip(r, 0.5, 2.5)
    k = np.clip(np.searchsorted(nodes, rc, side="right") - 1, 0, 15)
    want_f = rc * rc + (rc - nodes[k]) * (nodes[k + 1] - rc)
    assert np.max(np.abs(f - want_f)) < 1e-13
    assert np.max(np.abs(e - (3.0 - rc))) < 1e-13
    # at the nodes the interpolant is the tabulated value
    fn, en = orc.table_eval(0.5, 2.5, rows, nodes)
    assert np.array_equal(fn, rows[:, 2].astype(np.float64)) and np.array_equal(en, rows[:, 1].astype(np.float64))


def _spaced_points(n, L, dmin, seed):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        c = rng.uniform(0, L, 3)
        if pts:
            d = np.array(pts) - c
            d -= L * np.rint(d / L)
            if np.min(np.linalg.norm(d, axis=1)) < dmin:
                continue
        pts.append(c)
    return np.array(pts, F)


def test_table_pair_sums_brute_force(orc):
    """Pair sums with a table on the (0, 1) pair only, against an O(N^2) loop
    over minimum images in FP64 using the interpolant: force on i is
    (f/r)(r) d_ij, half the pair energy and half the pair virial per atom."""
    L, rc = 6.0, 2.5
    xyz = _spaced_points(48, L, 0.8, 3)
    types = np.arange(48) % 2
    p4 = pos4_of(xyz, types)
    rs = np.linspace(0.7, rc, 300)
    rows = np.stack([rs, 2.0 * np.exp(-rs), np.sin(3 * rs) / rs], axis=1).astype(F)
    tables = ([-1, 0, 0, -1], [(0.7, rc, rows)])
    lj0 = np.zeros(8, F)
    f8, vir = orc.forces_f64(p4, F(L), lj0, ntypes=2, rc=rc, terms=1, tables=tables)
    x = xyz.astype(np.float64)
    fb = np.zeros((48, 4))
    vb = np.zeros(6)
    for i in range(48):
        for j in range(48):
            if i == j or types[i] == types[j]:
                continue
            dv = x[i] - x[j]
            dv -= L * np.rint(dv / L)
            r = np.linalg.norm(dv)
            if r >= rc:
                continue
            fr, e = orc.table_eval(0.7, rc, rows, [r])
            fb[i, :3] += fr[0] * dv
            fb[i, 3] += 0.5 * e[0]
            vb += 0.5 * fr[0] * np.array([dv[0] * dv[0], dv[0] * dv[1], dv[0] * dv[2],
                                          dv[1] * dv[1], dv[1] * dv[2], dv[2] * dv[2]])
    assert np.abs(fb).max() > 0.1
    assert np.max(np.abs(f8 - fb)) < 1e-10 * np.abs(fb).max()
    assert np.max(np.abs(vir - vb)) < 1e-10 * np.abs(vb).max()
    # momentum conservation of a pair force
    assert np.max(np.abs(f8[:, :3].sum(axis=0))) < 1e-9


def lj_table_rows(n=10000, rlo=0.01, rhi=2.5, eps=1.0, sig=1.0):
    """An LJ table with the force/r clamp of src/force/lj.cu:48 (the construction
    of tests/test_table.cu:29-45)."""
    r = rlo + (rhi - rlo) / (n - 1) * np.arange(n)
    sr6 = (sig / r) ** 6
    fr = np.clip(24.0 * eps / (r * r) * (2.0 * sr6 * sr6 - sr6), -1000.0, 1000.0)
    return np.stack([r, 4.0 * eps * (sr6 * sr6 - sr6), fr], axis=1).astype(F)


def test_lj_table_matches_lj(orc):
    """tests/test_table.cu:8-115: an LJ table of 10,000 points reproduces the
    LJ forces and energies within the interpolation tolerance (1e-3)."""
    L = 7.0
    xyz, _ = _saw_chain(60, L, seed=4)
    p4 = pos4_of(xyz)
    fl, vl = orc.forces_f64(p4, F(L), KG_LJ, terms=1)
    ft, vt = orc.forces_f64(p4, F(L), np.zeros(2, F), terms=1, tables=([0], [(0.01, 2.5, lj_table_rows())]))
    floor = np.sqrt(np.mean(np.sum(fl[:, :3] ** 2, axis=1)))
    err = np.max(np.linalg.norm(ft[:, :3] - fl[:, :3], axis=1) / np.maximum(np.linalg.norm(fl[:, :3], axis=1), floor))
    assert 0 < err < 1e-3
    assert abs(ft[:, 3].sum() - fl[:, 3].sum()) < 1e-3 * abs(fl[:, 3].sum())
    assert np.max(np.abs(vt - vl)) < 1e-2 * np.abs(vl).max()


# ---------------------------------------------------------------------------
# Multiple-tau stress correlator (src/analysis/correlator.cu:33-92)
# ---------------------------------------------------------------------------
def test_corr_constant_signal(orc):
    # tests/test_correlator.cu:6-38: 64 pushes of (1, 0, ...) -> C_xx = 1 at every lag
    c, n = orc.corr_run(np.tile([1.0, 0, 0, 0, 0, 0], (64, 1)))
    seen = n[0] > 0
    assert seen.all()
    assert np.allclose(c[0, seen, 0] / n[0, seen], 1.0, atol=1e-12)
    assert not c[:, :, 1:].any()


def test_corr_decaying_signal(orc):
    # tests/test_correlator.cu:40-72: exp(-i / 10) decays with lag
    x = np.zeros((256, 6))
    x[:, 0] = np.exp(-np.arange(256, dtype=np.float32) / np.float32(10.0))
    c, n = orc.corr_run(x)
    assert c[0, 0, 0] / n[0, 0] > c[0, 1, 0] / n[0, 1]


def corr_by_levels(x):
    """The correlator restated over whole sequences: level l + 1's samples are
    means of consecutive blocks of 16 of level l's; at level l, sample t pairs
    with sample t - lag for every lag <= t mod 16."""
    c = np.zeros((16, 16, 6))
    n = np.zeros((16, 16), np.int64)
    seq = np.asarray(x, np.float64)
    for lv in range(16):
        t = np.arange(len(seq))
        for lag in range(16):
            sel = t[(t % 16) >= lag]
            c[lv, lag] = (seq[sel] * seq[sel - lag]).sum(axis=0)
            n[lv, lag] = len(sel)
        m = len(seq) // 16
        if m == 0:
            break
        seq = seq[:16 * m].reshape(m, 16, 6).mean(axis=1)
    return c, n


@pytest.mark.parametrize("nsamples", [1, 15, 16, 17, 300, 5000])
def test_corr_equals_level_sequence_form(orc, nsamples):
    x = np.random.default_rng(nsamples).normal(size=(nsamples, 6))
    c, n = orc.corr_run(x)
    cw, nw = corr_by_levels(x)
    assert np.array_equal(n, nw)
    assert np.allclose(c, cw, rtol=1e-12, atol=1e-12 * np.abs(cw).max())
    # level-0 counts in closed form: pushes t < N with t mod 16 >= lag
    lag = np.arange(16)
    assert np.array_equal(n[0], (nsamples // 16) * (16 - lag) + np.maximum(0, nsamples % 16 - lag))
