"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as a compute path of the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class OrcSys(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("ntypes", C.c_int), ("pos4", _fp), ("vel3", _fp), ("img3", _ip),
        ("masses", _fp), ("lj2", _fp), ("nbonds", C.c_int), ("bonds", _ip), ("btype", _ip),
        ("bp", _fp), ("nang", C.c_int), ("ang4", _ip), ("ap", _fp),
        ("L", C.c_float), ("rc", C.c_float), ("skin", C.c_float), ("dt", C.c_float),
        ("ensemble", C.c_int), ("T_start", C.c_float), ("T_stop", C.c_float), ("Tdamp", C.c_float),
        ("P_start", C.c_float), ("P_stop", C.c_float), ("Pdamp", C.c_float), ("chain", C.c_int),
        ("seed", C.c_uint64), ("nsteps_total", C.c_long), ("clamp_lj", C.c_int),
        ("clamp_fene", C.c_int), ("fcap", C.c_float),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_min_image.restype = C.c_float
        L.orc_min_image.argtypes = [C.c_float] * 3
        L.orc_min_image_fixed.restype = C.c_float
        L.orc_min_image_fixed.argtypes = [C.c_float] * 3
        L.orc_lj_inside_grid.restype = None
        L.orc_lj_inside_grid.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_void_p]
        L.orc_wrap.restype = C.c_float
        L.orc_wrap.argtypes = [C.c_float] * 3
        L.orc_morton3d.restype = C.c_uint32
        L.orc_morton3d.argtypes = [C.c_float] * 4
        L.orc_morton_perm.argtypes = [_fp, C.c_int, C.c_float, _ip]
        L.orc_cell_ids.restype = C.c_int
        L.orc_cell_ids.argtypes = [_fp, C.c_int, C.c_float, C.c_float, _ip]
        L.orc_build_exclusions.restype = C.c_int
        L.orc_build_exclusions.argtypes = [C.c_int, _ip, C.c_int, _ip, _ip, C.c_int]
        L.orc_pair_set.restype = C.c_long
        L.orc_pair_set.argtypes = [_fp, C.c_int, C.c_float, C.c_float, _ip, _ip, C.c_int, _ip, C.c_long]
        L.orc_lj_literal.argtypes = [_fp, C.c_int, _fp, C.c_int, C.c_float, C.c_float, _ip, _ip, _fp, _fp]
        L.orc_fene_literal.argtypes = [_fp, C.c_int, _ip, _ip, C.c_int, _fp, C.c_float, _fp, _fp]
        L.orc_angle_literal.argtypes = [_fp, C.c_int, _ip, C.c_int, _fp, C.c_float, _fp, _fp]
        L.orc_forces_f64.argtypes = [_fp, C.c_int, C.c_float, C.c_int, _fp, C.c_float, _fp, _fp,
                                     C.c_int, C.c_int, C.c_float, C.c_int, _ip, _ip, _ip, _ip,
                                     C.c_int, _ip, C.c_int, _dp, _dp]
        L.orc_forces_tab_f64.argtypes = [_fp, C.c_int, C.c_float, C.c_int, _fp, C.c_float, _fp, _fp,
                                         C.c_int, C.c_int, C.c_float, C.c_int, _ip, _ip, _ip, _ip,
                                         C.c_int, _ip, C.c_int, C.c_int, _ip, _ip, _fp, _fp, _dp, _dp]
        L.orc_table_eval.argtypes = [C.c_int, C.c_float, C.c_float, _fp, C.c_int, _dp, _dp, _dp]
        L.orc_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_normal3.argtypes = [C.c_uint64, C.c_uint32, C.c_int64, _fp]
        L.orc_nh_chain.restype = C.c_double
        L.orc_nh_chain.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_double,
                                   C.c_double, C.c_int, C.c_double]
        L.orc_md_run.restype = C.c_int
        L.orc_md_run.argtypes = [C.POINTER(OrcSys), C.c_long, _dp, _dp]
        L.orc_lj_cut_energy.restype = C.c_double
        L.orc_lj_cut_energy.argtypes = [_fp, C.c_int, C.c_int, _fp, C.c_float, C.c_float, _ip, C.c_int]
        L.orc_nh_chain_energy.restype = C.c_double
        L.orc_nh_chain_energy.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_double,
                                          C.c_int]
        L.orc_thermo.argtypes = [C.c_int, _fp, _fp, _fp, _dp, _dp, C.c_float, _dp]
        L.orc_cpu_md_langevin.restype = C.c_long
        L.orc_cpu_md_langevin.argtypes = [C.c_int, _fp, _fp, _ip, _fp, C.c_float, C.c_float, C.c_float,
                                          C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                          C.c_uint64, C.c_int, _ip, _fp, C.c_int, _ip, _fp, C.c_long,
                                          C.c_int]
        L.orc_cpu_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def type_bits(t):
    return np.array(t, dtype=np.int32).view(np.float32)


def min_image(dr, L):
    return lib().orc_min_image(dr, L, np.float32(1.0) / np.float32(L))


def lj_inside_grid(a, b, rc2, L):
    """Device LJ cutoff decisions (grid units) for point pairs a[k], b[k]."""
    a = np.ascontiguousarray(a, np.float32).reshape(-1, 3)
    b = np.ascontiguousarray(b, np.float32).reshape(-1, 3)
    out = np.zeros(len(a), np.int32)
    lib().orc_lj_inside_grid(len(a), a.ctypes.data, b.ctypes.data, rc2, L, out.ctypes.data)
    return out.astype(bool)


def wrap(x, L):
    return lib().orc_wrap(x, L, np.float32(1.0) / np.float32(L))


def morton3d(x, y, z, L):
    return lib().orc_morton3d(x, y, z, np.float32(1.0) / np.float32(L))


def morton_perm(pos4, L):
    pos4 = f32(pos4)
    n = pos4.shape[0]
    out = np.zeros(n, np.int32)
    lib().orc_morton_perm(_p(pos4, C.c_float), n, np.float32(1.0) / np.float32(L), _p(out, C.c_int32))
    return out


def cell_ids(pos4, L, rcs):
    pos4 = f32(pos4)
    out = np.zeros(pos4.shape[0], np.int32)
    nx = lib().orc_cell_ids(_p(pos4, C.c_float), pos4.shape[0], L, rcs, _p(out, C.c_int32))
    return out, nx


def exclusions(n, bonds):
    bonds = i32(bonds).reshape(-1, 2)
    off = np.zeros(n + 1, np.int32)
    cap = 2 * len(bonds) + 1
    lst = np.zeros(cap, np.int32)
    m = lib().orc_build_exclusions(n, _p(bonds, C.c_int32), len(bonds), _p(off, C.c_int32),
                                   _p(lst, C.c_int32), cap)
    return off, lst[:m]


def pair_set(pos4, L, rcs, excl=None, method=0):
    pos4 = f32(pos4)
    n = pos4.shape[0]
    eo, el = (None, None) if excl is None else (i32(excl[0]), i32(excl[1]))
    args = (_p(pos4, C.c_float), n, L, rcs, _p(eo, C.c_int32), _p(el, C.c_int32), method)
    np_ = lib().orc_pair_set(*args, None, 0)
    out = np.zeros((max(1, np_), 2), np.int32)
    lib().orc_pair_set(*args, _p(out, C.c_int32), np_)
    return out[:np_]


def lj_literal(pos4, lj, ntypes, rc, L, excl=None):
    pos4 = f32(pos4)
    n = pos4.shape[0]
    f = np.zeros((n, 4), np.float32)
    v = np.zeros(6, np.float32)
    eo, el = (None, None) if excl is None else (i32(excl[0]), i32(excl[1]))
    lib().orc_lj_literal(_p(pos4, C.c_float), n, _p(f32(lj), C.c_float), ntypes, rc, L,
                         _p(eo, C.c_int32), _p(el, C.c_int32), _p(f, C.c_float), _p(v, C.c_float))
    return f, v


def fene_literal(pos4, bonds, btypes, bp, L):
    pos4, bonds, btypes, bp = f32(pos4), i32(bonds), i32(btypes), f32(bp)
    n = pos4.shape[0]
    f = np.zeros((n, 4), np.float32)
    v = np.zeros(6, np.float32)
    lib().orc_fene_literal(_p(pos4, C.c_float), n, _p(bonds, C.c_int32), _p(btypes, C.c_int32),
                           len(btypes), _p(bp, C.c_float), L, _p(f, C.c_float), _p(v, C.c_float))
    return f, v


def angle_literal(pos4, ang4, ap, L):
    pos4, ang4, ap = f32(pos4), i32(ang4).reshape(-1, 4), f32(ap)
    n = pos4.shape[0]
    f = np.zeros((n, 4), np.float32)
    v = np.zeros(6, np.float32)
    lib().orc_angle_literal(_p(pos4, C.c_float), n, _p(ang4, C.c_int32), len(ang4), _p(ap, C.c_float),
                            L, _p(f, C.c_float), _p(v, C.c_float))
    return f, v


def pack_tables(tables):
    """(table_idx, [(rmin, rmax, rows (n, 3) of r, energy, force/r), ...]) as
    flat arrays: idx, points per table, ranges, rows."""
    tidx, tabs = tables
    tn = i32([np.asarray(t[2]).shape[0] for t in tabs])
    tr = f32([[t[0], t[1]] for t in tabs]).reshape(-1)
    td = f32(np.concatenate([np.asarray(t[2], np.float32).reshape(-1, 3) for t in tabs]))
    return i32(tidx).reshape(-1), tn, tr, td


def table_eval(rmin, rmax, rows, r):
    """FP64 linear interpolation of one table at distances r: (force/r, energy)."""
    rows = f32(rows).reshape(-1, 3)
    r = np.ascontiguousarray(r, np.float64).reshape(-1)
    f = np.zeros_like(r)
    e = np.zeros_like(r)
    lib().orc_table_eval(rows.shape[0], rmin, rmax, _p(rows, C.c_float), len(r), _p(r, C.c_double),
                         _p(f, C.c_double), _p(e, C.c_double))
    return f, e


def forces_f64(pos4, L, lj, ntypes=1, rc=2.5, bonds=None, btypes=None, bp=None, ang4=None, ap=None,
               terms=7, clamp_lj=1, clamp_fene=1, fcap=0.0, tables=None):
    pos4 = f32(pos4)
    n = pos4.shape[0]
    bonds = i32(np.zeros((0, 2)) if bonds is None else bonds).reshape(-1, 2)
    btypes = i32(np.zeros(len(bonds)) if btypes is None else btypes)
    ang4 = i32(np.zeros((0, 4)) if ang4 is None else ang4).reshape(-1, 4)
    bp = f32(np.zeros(4) if bp is None else bp)
    ap = f32(np.zeros(2) if ap is None else ap)
    eo, el = exclusions(n, bonds)
    f8 = np.zeros((n, 4), np.float64)
    vir = np.zeros(6, np.float64)
    if tables is not None:
        tidx, tn, tr, td = pack_tables(tables)
        lib().orc_forces_tab_f64(_p(pos4, C.c_float), n, L, ntypes, _p(f32(lj), C.c_float), rc,
                                 _p(bp, C.c_float), _p(ap, C.c_float), clamp_lj, clamp_fene, fcap, terms,
                                 _p(eo, C.c_int32), _p(el, C.c_int32), _p(bonds, C.c_int32),
                                 _p(btypes, C.c_int32), len(bonds), _p(ang4, C.c_int32), len(ang4),
                                 len(tn), _p(tidx, C.c_int32), _p(tn, C.c_int32), _p(tr, C.c_float),
                                 _p(td, C.c_float), _p(f8, C.c_double), _p(vir, C.c_double))
        return f8, vir
    lib().orc_forces_f64(_p(pos4, C.c_float), n, L, ntypes, _p(f32(lj), C.c_float), rc,
                         _p(bp, C.c_float), _p(ap, C.c_float), clamp_lj, clamp_fene, fcap, terms,
                         _p(eo, C.c_int32), _p(el, C.c_int32), _p(bonds, C.c_int32),
                         _p(btypes, C.c_int32), len(bonds), _p(ang4, C.c_int32), len(ang4),
                         _p(f8, C.c_double), _p(vir, C.c_double))
    return f8, vir


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def normal3(seed, gid, step):
    z = np.zeros(3, np.float32)
    lib().orc_normal3(seed, gid, step, _p(z, C.c_float))
    return z


def nh_chain(prec, M, Q1, Qr, xi, vxi, KE, kT, natoms, half_dt):
    xi = np.array(xi, np.float64)
    vxi = np.array(vxi, np.float64)
    s = lib().orc_nh_chain(prec, M, Q1, Qr, _p(xi, C.c_double), _p(vxi, C.c_double), KE, kT, natoms,
                           half_dt)
    return s, xi, vxi


def rg(pos4, image, mol, L):
    """Chain-averaged radius of gyration, FP64 (chains = runs of equal mol id)."""
    p = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
    im = np.ascontiguousarray(image, np.int32).reshape(-1, 3)
    m = np.ascontiguousarray(mol, np.int32).reshape(-1)
    f = lib().orc_rg
    f.restype = C.c_double
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
    return f(len(p), p.ctypes.data, im.ctypes.data, m.ctypes.data, float(L))


def corr_run(samples):
    """A fresh multiple-tau correlator fed the samples (n, 6): sums (16, 16, 6), counts (16, 16)."""
    x = np.ascontiguousarray(samples, np.float64).reshape(-1, 6)
    c = np.zeros((16, 16, 6), np.float64)
    n = np.zeros((16, 16), np.int32)
    f = lib().orc_corr_run
    f.restype = None
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    f(len(x), x.ctypes.data, c.ctypes.data, n.ctypes.data)
    return c, n


EXT = 40  # doubles per extended-system record of md_run(..., ext=True)


def md_run(sysd, nsteps, ext=False):
    """Runs the oracle MD driver on a dict describing the system (see tests).
    Returns (pos4, vel3, img3, L, thermo[nsteps, 11]) and, with ext=True, the
    extended-system records [nsteps + 1, EXT] (row 0 = the initial state; the
    column meanings are listed at orc_md_run in oracle_capi.cpp)."""
    n = sysd["pos4"].shape[0]
    keep = {k: (f32(v) if v.dtype.kind == "f" else i32(v)) for k, v in sysd.items() if isinstance(v, np.ndarray)}
    keep["pos4"] = keep["pos4"].copy()
    keep["vel3"] = keep.get("vel3", np.zeros((n, 3), np.float32)).copy()
    keep["img3"] = keep.get("img3", np.zeros((n, 3), np.int32)).copy()
    for k, shape, dt in (("bonds", (0, 2), np.int32), ("btype", (0,), np.int32), ("ang4", (0, 4), np.int32),
                         ("bp", (1, 4), np.float32), ("ap", (1, 2), np.float32)):
        keep.setdefault(k, np.zeros(shape, dt))
    s = OrcSys()
    s.n, s.ntypes = n, sysd.get("ntypes", 1)
    s.pos4, s.vel3, s.img3 = _p(keep["pos4"], C.c_float), _p(keep["vel3"], C.c_float), _p(keep["img3"], C.c_int32)
    s.masses, s.lj2 = _p(keep["masses"], C.c_float), _p(keep["lj"], C.c_float)
    s.nbonds, s.bonds, s.btype, s.bp = len(keep["btype"]), _p(keep["bonds"], C.c_int32), _p(keep["btype"], C.c_int32), _p(keep["bp"], C.c_float)
    s.nang, s.ang4, s.ap = len(keep["ang4"]), _p(keep["ang4"], C.c_int32), _p(keep["ap"], C.c_float)
    s.L, s.rc, s.skin, s.dt = sysd["L"], sysd.get("rc", 2.5), sysd.get("skin", 0.3), sysd.get("dt", 0.001)
    s.ensemble = sysd.get("ensemble", 0)
    s.T_start, s.T_stop, s.Tdamp = sysd.get("T_start", 1.0), sysd.get("T_stop", 1.0), sysd.get("Tdamp", 1.0)
    s.P_start, s.P_stop, s.Pdamp = sysd.get("P_start", 0.0), sysd.get("P_stop", 0.0), sysd.get("Pdamp", 1.0)
    s.chain, s.seed, s.nsteps_total = sysd.get("chain", 3), sysd.get("seed", 42), sysd.get("nsteps_total", nsteps)
    s.clamp_lj, s.clamp_fene, s.fcap = 1, 1, sysd.get("fcap", 0.0)
    thermo = np.zeros((nsteps, 11), np.float64)
    xr = np.zeros((nsteps + 1, EXT), np.float64) if ext else None
    lib().orc_md_run(C.byref(s), nsteps, _p(thermo, C.c_double), _p(xr, C.c_double) if ext else None)
    if ext:
        return keep["pos4"], keep["vel3"], keep["img3"], float(s.L), thermo, xr
    return keep["pos4"], keep["vel3"], keep["img3"], float(s.L), thermo


def lj_cut_energy(pos4, L, lj, bonds=None, ntypes=1, rc=2.5):
    """Sum of the LJ energy at the cutoff over the non-excluded pairs inside
    it: U - lj_cut_energy is the energy of the cut-and-shifted potential."""
    pos4 = f32(pos4).reshape(-1, 4)
    b = i32(np.zeros((0, 2)) if bonds is None else bonds).reshape(-1, 2)
    lj = f32(lj)
    return lib().orc_lj_cut_energy(_p(pos4, C.c_float), len(pos4), ntypes, _p(lj, C.c_float), rc, L,
                                   _p(b, C.c_int32), len(b))


def nh_chain_energy(form, M, Q1, Qr, xi, vxi, kT, natoms):
    """Chain energy: form 0 = the chain terms of the reference's conserved
    quantity (tests/test_nose_hoover.cu:19-40); form 1 = the weights the
    chain's dynamics conserves (oracle.hpp nh_chain_energy_dyn)."""
    xi = np.ascontiguousarray(np.resize(np.asarray(xi, np.float64), 10))
    vxi = np.ascontiguousarray(np.resize(np.asarray(vxi, np.float64), 10))
    return lib().orc_nh_chain_energy(form, int(M), float(Q1), float(Qr), _p(xi, C.c_double), _p(vxi, C.c_double),
                                     float(kT), int(natoms))


def extended_energy(xr, natoms, dt, form=1, shifted=True):
    """Extended-system energy of each md_run ext record:
    K + U + chain energy (form as in nh_chain_energy), and under NPT
    + P V + v_eps^2 / (2 W) with V the barostat's volume.  The reference's
    barostat scales velocities by exp(-(1 + 3/Nf) v_eps/W dt/2) but its
    v_eps force lacks the matching (3/Nf) 2K term of the MTK equations
    (src/integrate/nose_hoover.cu:206-222,236-238), so under NPT the
    energy changes at the rate -(6K/Nf) v_eps/W; the NPT value adds the
    trapezoid integral of that rate, which makes it a constant of the
    continuous dynamics.  With shifted=True the potential energy is that of
    the cut-and-shifted LJ (same forces; the truncated LJ of
    src/force/lj.cu:37-50 jumps by U(rc) whenever a pair crosses the cutoff,
    which no integrator conserves).  Returns one value per record."""
    out = np.zeros(len(xr))
    Nf = 3.0 * natoms
    acc = 0.0
    for r, x in enumerate(xr):
        M = int(x[9])
        e = x[0] + x[1] - (x[32] if shifted else 0.0) + nh_chain_energy(form, M, x[4], x[5], x[10:20],
                                                                         x[20:30], x[30], natoms)
        if x[6] > 0.0:
            if r > 0:
                p = xr[r - 1]
                acc += 0.5 * dt * ((6.0 * p[0] / Nf) * p[8] / p[6] + (6.0 * x[0] / Nf) * x[8] / x[6])
            e += x[31] * x[3] + 0.5 * x[8] * x[8] / x[6] + acc
        out[r] = e
    return out


def cpu_md_langevin(pos4, vel3, img3, masses, L, nsteps, bonds, bp, ang4, ap, T=1.0, Tdamp=1.0, seed=42,
                    eps=1.0, sig=1.0, rc=2.5, skin=0.3, dt=0.001, threads=0):
    pos4, vel3, img3 = f32(pos4).copy(), f32(vel3).copy(), i32(img3).copy()
    bonds, ang4 = i32(bonds).reshape(-1, 2), i32(ang4).reshape(-1, 4)
    bp, ap, masses = f32(bp).reshape(-1), f32(ap).reshape(-1), f32(masses)
    nreb = lib().orc_cpu_md_langevin(pos4.shape[0], _p(pos4, C.c_float), _p(vel3, C.c_float),
                                     _p(img3, C.c_int32), _p(masses, C.c_float), eps, sig, rc, skin, L, dt,
                                     T, Tdamp, seed, len(bonds), _p(bonds, C.c_int32), _p(bp, C.c_float),
                                     len(ang4), _p(ang4, C.c_int32), _p(ap, C.c_float), nsteps, threads)
    return pos4, vel3, img3, nreb


def literal_md_langevin(pos4, vel3, img3, masses, L, nsteps, lj, bonds, btypes, bp, ang4, ap, T=1.0, Tdamp=1.0,
                        seed=42, rc=2.5, dt=0.001, ntypes=1):
    """The literal brute-force CPU path (one core): see orc_literal_md_langevin."""
    pos4, vel3, img3 = f32(pos4).copy(), f32(vel3).copy(), i32(img3).copy()
    bonds, ang4 = i32(bonds).reshape(-1, 2), i32(ang4).reshape(-1, 4)
    btypes, bp, ap, masses, lj = i32(btypes), f32(bp).reshape(-1), f32(ap).reshape(-1), f32(masses), f32(lj)
    f = lib().orc_literal_md_langevin
    f.restype = None
    f.argtypes = [C.c_int, _fp, _fp, _ip, _fp, _fp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                  C.c_float, C.c_uint64, C.c_int, _ip, _ip, _fp, C.c_int, _ip, _fp, C.c_long]
    f(pos4.shape[0], _p(pos4, C.c_float), _p(vel3, C.c_float), _p(img3, C.c_int32), _p(masses, C.c_float),
      _p(lj, C.c_float), ntypes, rc, L, dt, T, Tdamp, seed, len(bonds), _p(bonds, C.c_int32), _p(btypes, C.c_int32),
      _p(bp, C.c_float), len(ang4), _p(ang4, C.c_int32), _p(ap, C.c_float), nsteps)
    return pos4, vel3, img3


def cpu_threads():
    return lib().orc_cpu_threads()
