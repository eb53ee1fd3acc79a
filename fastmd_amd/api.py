"""ctypes binding of libfastmd.so (include/fmd/fmd.h).

The Python side is a thin host mirror of the C ABI: every call goes to the
native library, which owns all device memory.  There is no Python or CPU
compute path; if the library is missing, loading raises.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libfastmd.so")

OK = 0
ERR_ARG, ERR_IO, ERR_PARSE, ERR_HIP, ERR_NOMEM, ERR_BUDGET = 1, 2, 3, 4, 5, 6
CORR_LEVELS, CORR_POINTS = 16, 16
ERR_NBR_OVERFLOW, ERR_STRETCH, ERR_NAN, ERR_STATE, ERR_CELL = 7, 8, 9, 10, 11
LANGEVIN, NVT_NH, NPT_NH = 0, 1, 2
TERM_LJ, TERM_BOND, TERM_ANGLE, TERM_ALL = 1, 2, 4, 7
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)


class SystemDesc(C.Structure):
    _fields_ = [
        ("natoms", C.c_int32), ("ntypes", C.c_int32),
        ("pos", _fp), ("vel", _fp), ("image", _ip), ("mol_id", _ip),
        ("nbonds", C.c_int32), ("bonds", _ip), ("bond_types", _ip),
        ("nangles", C.c_int32), ("angles", _ip),
        ("masses", _fp), ("lj", _fp),
        ("nbond_types", C.c_int32), ("bond_params", _fp),
        ("nangle_types", C.c_int32), ("angle_params", _fp),
        ("L", C.c_float), ("rc", C.c_float), ("skin", C.c_float), ("dt", C.c_float),
        ("ensemble", C.c_int32),
        ("T_start", C.c_float), ("T_stop", C.c_float), ("Tdamp", C.c_float),
        ("P_start", C.c_float), ("P_stop", C.c_float), ("Pdamp", C.c_float),
        ("chain_length", C.c_int32), ("seed", C.c_uint64), ("nsteps", C.c_int64),
        ("thermo_freq", C.c_int32), ("force_cap", C.c_float),
        ("ntables", C.c_int32), ("table_idx", _ip), ("table_n", _ip), ("table_range", _fp),
        ("table_data", _fp), ("stress_freq", C.c_int32),
    ]


class Thermo(C.Structure):
    _fields_ = [("step", C.c_int64), ("ke", C.c_double), ("pe", C.c_double),
                ("temp", C.c_double), ("press", C.c_double * 6), ("L", C.c_double),
                ("max_v", C.c_double), ("max_f", C.c_double), ("max_v_id", C.c_int64), ("max_f_id", C.c_int64)]

    def as_dict(self):
        return {"step": self.step, "ke": self.ke, "pe": self.pe, "temp": self.temp,
                "press": list(self.press), "L": self.L, "max_v": self.max_v, "max_f": self.max_f,
                "max_v_id": self.max_v_id, "max_f_id": self.max_f_id}


MAX_CHAIN = 10


class Thermostat(C.Structure):
    _fields_ = [("ensemble", C.c_int32), ("M", C.c_int32), ("step", C.c_int64),
                ("Q1", C.c_double), ("Qr", C.c_double), ("W", C.c_double), ("V0", C.c_double),
                ("eps", C.c_double), ("veps", C.c_double),
                ("xi", C.c_double * MAX_CHAIN), ("vxi", C.c_double * MAX_CHAIN), ("ke_carry", C.c_double)]

    def as_dict(self):
        return {"ensemble": self.ensemble, "M": self.M, "step": self.step, "Q1": self.Q1, "Qr": self.Qr,
                "W": self.W, "V0": self.V0, "eps": self.eps, "veps": self.veps,
                "xi": np.array(self.xi[:]), "vxi": np.array(self.vxi[:]), "ke_carry": self.ke_carry}


class ChainRun(C.Structure):
    _fields_ = [("M", C.c_int32), ("natoms", C.c_int32), ("carry", C.c_int32), ("n", C.c_int32),
                ("Q1", C.c_double), ("Qr", C.c_double), ("kT", C.c_double), ("half_dt", C.c_double),
                ("ke", _dp)]


class ConfExtra(C.Structure):
    _fields_ = [("thermo_on", C.c_int32), ("thermo_freq", C.c_int32), ("thermo_file", C.c_char * 256),
                ("restart_freq", C.c_int32), ("restart_file", C.c_char * 256),
                ("forces_dump_file", C.c_char * 256), ("dump_freq", C.c_int32),
                ("rg_on", C.c_int32), ("rg_freq", C.c_int32), ("rg_file", C.c_char * 256),
                ("stress_on", C.c_int32), ("stress_freq", C.c_int32), ("stress_file", C.c_char * 256)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_jobs", C.c_int32), ("graph_steps", C.c_int32)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("natoms", C.c_int32), ("steps", C.c_int64),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("last", Thermo)]


_lib = None

# (name, restype, argtypes) of every exported entry point of include/fmd/fmd.h
SIGNATURES = [
    ("fmd_version", C.c_char_p, []),
    ("fmd_last_host_error", C.c_char_p, []),
    ("fmd_topo_load_conf", C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("fmd_topo_load_data", C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("fmd_topo_kg_melt", C.c_int32, [C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("fmd_topo_desc", C.c_int32, [C.c_void_p, C.POINTER(SystemDesc)]),
    ("fmd_topo_conf_extra", C.c_int32, [C.c_void_p, C.POINTER(ConfExtra)]),
    ("fmd_topo_exclusions", C.c_int64, [C.c_void_p, _ip, _ip, C.c_int64]),
    ("fmd_topo_write_data", C.c_int32, [C.c_void_p, C.c_char_p, _fp, _fp, _ip, C.c_float, C.c_int64]),
    ("fmd_topo_free", None, [C.c_void_p]),
    ("fmd_engine_create", C.c_int32, [C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("fmd_engine_destroy", None, [C.c_void_p]),
    ("fmd_last_error", C.c_char_p, [C.c_void_p]),
    ("fmd_system_bytes", C.c_uint64, [C.POINTER(SystemDesc)]),
    ("fmd_engine_bytes_in_use", C.c_uint64, [C.c_void_p]),
    ("fmd_system_create", C.c_int32, [C.c_void_p, C.POINTER(SystemDesc), _ip]),
    ("fmd_system_destroy", C.c_int32, [C.c_void_p, C.c_int32]),
    ("fmd_system_set_force_cap", C.c_int32, [C.c_void_p, C.c_int32, C.c_float]),
    ("fmd_system_set_langevin", C.c_int32, [C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_float]),
    ("fmd_build_neighbors", C.c_int32, [C.c_void_p, C.c_int32]),
    ("fmd_get_pairs", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), _ip, C.c_int64]),
    ("fmd_get_cell_ids", C.c_int32, [C.c_void_p, C.c_int32, _ip, _ip]),
    ("fmd_get_morton_keys", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint32)]),
    ("fmd_compute_forces", C.c_int32, [C.c_void_p, C.c_int32, C.c_uint32, _fp, _dp, _dp]),
    ("fmd_step", C.c_int32, [C.c_void_p, _ip, C.c_int32, C.c_int64]),
    ("fmd_step_plain", C.c_int32, [C.c_void_p, C.c_int32, C.c_int64]),
    ("fmd_system_step", C.c_int64, [C.c_void_p, C.c_int32]),
    ("fmd_system_natoms", C.c_int32, [C.c_void_p, C.c_int32]),
    ("fmd_system_error", C.c_int32, [C.c_void_p, C.c_int32]),
    ("fmd_get_state", C.c_int32, [C.c_void_p, C.c_int32, _fp, _fp, _ip]),
    ("fmd_get_box", C.c_int32, [C.c_void_p, C.c_int32, _dp]),
    ("fmd_get_state_box", C.c_int32, [C.c_void_p, C.c_int32, _dp]),
    ("fmd_get_rg", C.c_int32, [C.c_void_p, C.c_int32, _dp, _dp, C.c_int32, _ip]),
    ("fmd_get_rg_noalloc", C.c_int32, [C.c_void_p, C.c_int32, _dp, _dp, C.c_int32, _ip]),
    ("fmd_snapshot_size", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64)]),
    ("fmd_snapshot", C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64]),
    ("fmd_system_resume", C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, _ip]),
    ("fmd_corr_push", C.c_int32, [C.c_void_p, C.c_int32, _dp]),
    ("fmd_corr_reset", C.c_int32, [C.c_void_p, C.c_int32]),
    ("fmd_get_corr", C.c_int32, [C.c_void_p, C.c_int32, _dp, _ip, _ip]),
    ("fmd_get_thermo", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(Thermo)]),
    ("fmd_drain_thermo", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(Thermo), C.c_int32, _ip,
                                     C.POINTER(C.c_int64)]),
    ("fmd_get_thermostat", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(Thermostat)]),
    ("fmd_nh_chain_run", C.c_int32, [C.c_void_p, C.POINTER(ChainRun), _dp, _dp, _dp]),
    ("fmd_system_rebuilds", C.c_int64, [C.c_void_p, C.c_int32]),
    ("fmd_autotune", C.c_int32, [C.c_void_p, _ip, C.c_int32, C.c_int32]),
    ("fmd_tuned_block", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32]),
    ("fmd_autotune_times", C.c_int32, [C.c_void_p, C.c_int32, _ip, _ip, _dp, C.c_int32]),
    ("fmd_engine_set_tuned", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    ("fmd_time_force", C.c_int32, [C.c_void_p, _ip, C.c_int32, C.c_uint32, C.c_int32, _dp, _dp,
                                   C.POINTER(C.c_int64)]),
    ("fmd_size_class", C.c_int32, [C.c_int32]),
    ("fmd_server_create", C.c_int32, [C.c_int32, C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]),
    ("fmd_server_submit", C.c_int32, [C.c_void_p, C.POINTER(SystemDesc), C.c_int64, C.POINTER(C.c_int64)]),
    ("fmd_server_poll", C.c_int32, [C.c_void_p, C.c_int64, _ip]),
    ("fmd_server_result", C.c_int32, [C.c_void_p, C.c_int64, C.POINTER(JobResult), _fp, _fp]),
    ("fmd_server_drain", C.c_int32, [C.c_void_p]),
    ("fmd_server_now", C.c_double, [C.c_void_p]),
    ("fmd_server_stats", C.c_int32, [C.c_void_p, _dp, C.POINTER(C.c_int64), _ip, _ip]),
    ("fmd_server_destroy", None, [C.c_void_p]),
    ("fmd_server_set_block", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    ("fmd_admission_replay", C.c_int32, [C.c_int32, _dp, C.POINTER(C.c_uint64), _dp, C.c_uint64,
                                         C.c_int32, _dp, _dp]),
]


def lib():
    """The native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m fastmd_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class FmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fmd error {code}: {msg}")
        self.code = code


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _host_err():
    m = lib().fmd_last_host_error()
    return m.decode() if m else ""


class Desc:
    """A system descriptor that keeps its numpy arrays alive."""

    def __init__(self, pos4, L, lj, ntypes=1, vel=None, image=None, mol=None, bonds=None,
                 bond_types=None, bond_params=None, angles=None, angle_params=None, masses=None,
                 rc=2.5, skin=0.3, dt=0.001, ensemble=LANGEVIN, T=(1.0, 1.0), Tdamp=1.0,
                 P=(0.0, 0.0), Pdamp=1.0, chain=3, seed=42, nsteps=1000, thermo_freq=0,
                 force_cap=0.0, tables=None, stress_freq=0):
        """tables = (table_idx, [(rmin, rmax, rows), ...]): table_idx holds
        ntypes*ntypes entries (-1 = none), rows is (n, 3) of r, energy, force/r."""
        self.pos4 = np.ascontiguousarray(pos4, dtype=np.float32).reshape(-1, 4)
        n = self.pos4.shape[0]
        self.vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        self.image = None if image is None else np.ascontiguousarray(image, dtype=np.int32).reshape(n, 3)
        self.mol = None if mol is None else np.ascontiguousarray(mol, dtype=np.int32)
        self.bonds = None if bonds is None else np.ascontiguousarray(bonds, dtype=np.int32).reshape(-1, 2)
        nb = 0 if self.bonds is None else self.bonds.shape[0]
        if bond_types is None and nb:
            bond_types = np.zeros(nb, np.int32)
        self.bond_types = None if bond_types is None else np.ascontiguousarray(bond_types, dtype=np.int32)
        self.bond_params = None if bond_params is None else np.ascontiguousarray(bond_params, dtype=np.float32).reshape(-1, 4)
        self.angles = None if angles is None else np.ascontiguousarray(angles, dtype=np.int32).reshape(-1, 4)
        self.angle_params = None if angle_params is None else np.ascontiguousarray(angle_params, dtype=np.float32).reshape(-1, 2)
        self.masses = None if masses is None else np.ascontiguousarray(masses, dtype=np.float32)
        self.lj = np.ascontiguousarray(lj, dtype=np.float32).reshape(-1)
        d = SystemDesc()
        d.natoms, d.ntypes = n, ntypes
        d.pos, d.vel = _ptr(self.pos4, C.c_float), _ptr(self.vel, C.c_float)
        d.image, d.mol_id = _ptr(self.image, C.c_int32), _ptr(self.mol, C.c_int32)
        d.nbonds = nb
        d.bonds, d.bond_types = _ptr(self.bonds, C.c_int32), _ptr(self.bond_types, C.c_int32)
        d.nangles = 0 if self.angles is None else self.angles.shape[0]
        d.angles = _ptr(self.angles, C.c_int32)
        d.masses, d.lj = _ptr(self.masses, C.c_float), _ptr(self.lj, C.c_float)
        d.nbond_types = 0 if self.bond_params is None else self.bond_params.shape[0]
        d.bond_params = _ptr(self.bond_params, C.c_float)
        d.nangle_types = 0 if self.angle_params is None else self.angle_params.shape[0]
        d.angle_params = _ptr(self.angle_params, C.c_float)
        d.L, d.rc, d.skin, d.dt = L, rc, skin, dt
        d.ensemble = ensemble
        d.T_start, d.T_stop, d.Tdamp = T[0], T[1], Tdamp
        d.P_start, d.P_stop, d.Pdamp = P[0], P[1], Pdamp
        d.chain_length, d.seed, d.nsteps = chain, seed, nsteps
        d.thermo_freq, d.force_cap = thermo_freq, force_cap
        d.ntables = 0
        d.stress_freq = stress_freq
        if tables is not None:
            tidx, tabs = tables
            self.table_idx = np.ascontiguousarray(tidx, dtype=np.int32).reshape(-1)
            self.table_n = np.array([np.asarray(t[2]).shape[0] for t in tabs], np.int32)
            self.table_range = np.array([[t[0], t[1]] for t in tabs], np.float32).reshape(-1)
            self.table_data = np.ascontiguousarray(
                np.concatenate([np.asarray(t[2], np.float32).reshape(-1, 3) for t in tabs]), dtype=np.float32)
            d.ntables = len(tabs)
            d.table_idx, d.table_n = _ptr(self.table_idx, C.c_int32), _ptr(self.table_n, C.c_int32)
            d.table_range, d.table_data = _ptr(self.table_range, C.c_float), _ptr(self.table_data, C.c_float)
        self.c = d

    @property
    def natoms(self):
        return self.c.natoms


class Topo:
    """Host-side topology from run.conf / LAMMPS data / the KG generator."""

    def __init__(self, handle):
        self.h = handle
        self.c = SystemDesc()
        lib().fmd_topo_desc(self.h, C.byref(self.c))

    @classmethod
    def from_conf(cls, path):
        h = C.c_void_p()
        rc = lib().fmd_topo_load_conf(str(path).encode(), C.byref(h))
        if rc:
            raise FmdError(rc, _host_err())
        return cls(h)

    @classmethod
    def from_data(cls, path):
        h = C.c_void_p()
        rc = lib().fmd_topo_load_data(str(path).encode(), C.byref(h))
        if rc:
            raise FmdError(rc, _host_err())
        return cls(h)

    @classmethod
    def kg_melt(cls, nchains, chain_len, rho=0.85, seed=42):
        h = C.c_void_p()
        rc = lib().fmd_topo_kg_melt(nchains, chain_len, rho, seed, C.byref(h))
        if rc:
            raise FmdError(rc, _host_err())
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().fmd_topo_free(self.h)
            self.h = None

    def _arr(self, ptr, n, dtype, shape):
        if not ptr or n == 0:
            return np.zeros(shape, dtype)
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype).reshape(shape)

    @property
    def natoms(self):
        return self.c.natoms

    def arrays(self):
        d, n = self.c, self.c.natoms
        nti = d.ntypes * d.ntypes if d.ntables else 0
        tn = self._arr(d.table_n, d.ntables, np.int32, (d.ntables,))
        npts = int(tn.sum())
        return {
            "pos4": self._arr(d.pos, 4 * n, np.float32, (n, 4)),
            "vel": self._arr(d.vel, 3 * n, np.float32, (n, 3)),
            "image": self._arr(d.image, 3 * n, np.int32, (n, 3)),
            "mol": self._arr(d.mol_id, n, np.int32, (n,)),
            "bonds": self._arr(d.bonds, 2 * d.nbonds, np.int32, (d.nbonds, 2)),
            "bond_types": self._arr(d.bond_types, d.nbonds, np.int32, (d.nbonds,)),
            "angles": self._arr(d.angles, 4 * d.nangles, np.int32, (d.nangles, 4)),
            "masses": self._arr(d.masses, d.ntypes, np.float32, (d.ntypes,)),
            "lj": self._arr(d.lj, 2 * d.ntypes * d.ntypes, np.float32, (d.ntypes * d.ntypes, 2)),
            "bond_params": self._arr(d.bond_params, 4 * d.nbond_types, np.float32, (d.nbond_types, 4)),
            "angle_params": self._arr(d.angle_params, 2 * d.nangle_types, np.float32, (d.nangle_types, 2)),
            "table_idx": self._arr(d.table_idx, nti, np.int32, (nti,)),
            "table_n": tn,
            "table_range": self._arr(d.table_range, 2 * d.ntables, np.float32, (d.ntables, 2)),
            "table_data": self._arr(d.table_data, 3 * npts, np.float32, (npts, 3)),
        }

    def extra(self):
        e = ConfExtra()
        lib().fmd_topo_conf_extra(self.h, C.byref(e))
        return e

    def exclusions(self):
        n = self.natoms
        off = np.zeros(n + 1, np.int32)
        m = lib().fmd_topo_exclusions(self.h, _ptr(off, C.c_int32), None, 0)
        lst = np.zeros(max(1, m), np.int32)
        lib().fmd_topo_exclusions(self.h, _ptr(off, C.c_int32), _ptr(lst, C.c_int32), m)
        return off, lst[:m]

    def write_data(self, path, pos4=None, vel=None, image=None, L=0.0, step=0):
        rc = lib().fmd_topo_write_data(self.h, str(path).encode(), _ptr(pos4, C.c_float),
                                       _ptr(vel, C.c_float), _ptr(image, C.c_int32), L, step)
        if rc:
            raise FmdError(rc, _host_err())


class Engine:
    """One device-resident engine (one per device and host thread)."""

    def __init__(self, device=0, hbm_budget=0):
        h = C.c_void_p()
        rc = lib().fmd_engine_create(device, hbm_budget, C.byref(h))
        if rc:
            raise FmdError(rc, _host_err())
        self.h = h
        self.natoms = {}

    def close(self):
        if getattr(self, "h", None):
            lib().fmd_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc:
            raise FmdError(rc, lib().fmd_last_error(self.h).decode())

    def add(self, desc):
        """desc: Desc, Topo or SystemDesc -> system id."""
        c = desc.c if hasattr(desc, "c") else desc
        sid = C.c_int32()
        self._chk(lib().fmd_system_create(self.h, C.byref(c), C.byref(sid)))
        self.natoms[sid.value] = c.natoms
        return sid.value

    def remove(self, sid):
        self._chk(lib().fmd_system_destroy(self.h, sid))
        self.natoms.pop(sid, None)

    def set_force_cap(self, sid, cap):
        self._chk(lib().fmd_system_set_force_cap(self.h, sid, cap))

    def set_langevin(self, sid, T_start, T_stop, Tdamp):
        self._chk(lib().fmd_system_set_langevin(self.h, sid, T_start, T_stop, Tdamp))

    def build_neighbors(self, sid):
        self._chk(lib().fmd_build_neighbors(self.h, sid))

    def pairs(self, sid):
        n = C.c_int64()
        self._chk(lib().fmd_get_pairs(self.h, sid, C.byref(n), None, 0))
        ij = np.zeros((max(1, n.value), 2), np.int32)
        self._chk(lib().fmd_get_pairs(self.h, sid, C.byref(n), _ptr(ij, C.c_int32), n.value))
        return ij[:n.value]

    def cell_ids(self, sid):
        out = np.zeros(self.natoms[sid], np.int32)
        nx = C.c_int32()
        self._chk(lib().fmd_get_cell_ids(self.h, sid, _ptr(out, C.c_int32), C.byref(nx)))
        return out, nx.value

    def morton_keys(self, sid):
        out = np.zeros(self.natoms[sid], np.uint32)
        self._chk(lib().fmd_get_morton_keys(self.h, sid, _ptr(out, C.c_uint32)))
        return out

    def forces(self, sid, terms=TERM_ALL):
        f = np.zeros((self.natoms[sid], 4), np.float32)
        vir = np.zeros(6, np.float64)
        pe = C.c_double()
        self._chk(lib().fmd_compute_forces(self.h, sid, terms, _ptr(f, C.c_float), _ptr(vir, C.c_double),
                                           C.byref(pe)))
        return f, vir, pe.value

    def step(self, sids, nsteps):
        if isinstance(sids, int):
            sids = [sids]
        a = np.asarray(sids, np.int32)
        self._chk(lib().fmd_step(self.h, _ptr(a, C.c_int32), len(a), nsteps))

    def step_plain(self, sid, nsteps):
        self._chk(lib().fmd_step_plain(self.h, sid, nsteps))

    def steps_done(self, sid):
        return lib().fmd_system_step(self.h, sid)

    def error(self, sid):
        """The error code that stopped the system (OK while healthy)."""
        return lib().fmd_system_error(self.h, sid)

    def rebuilds(self, sid):
        return lib().fmd_system_rebuilds(self.h, sid)

    def state(self, sid):
        n = self.natoms[sid]
        p = np.zeros((n, 4), np.float32)
        v = np.zeros((n, 3), np.float32)
        im = np.zeros((n, 3), np.int32)
        self._chk(lib().fmd_get_state(self.h, sid, _ptr(p, C.c_float), _ptr(v, C.c_float), _ptr(im, C.c_int32)))
        return p, v, im

    def box(self, sid):
        L = C.c_double()
        self._chk(lib().fmd_get_box(self.h, sid, C.byref(L)))
        return L.value

    def rg(self, sid, per_chain=False):
        """Chain-averaged radius of gyration (and the per-chain values)."""
        rg, n = C.c_double(), C.c_int32()
        self._chk(lib().fmd_get_rg(self.h, sid, C.byref(rg), None, 0, C.byref(n)))
        if not per_chain:
            return rg.value
        out = np.zeros(n.value, np.float64)
        self._chk(lib().fmd_get_rg(self.h, sid, C.byref(rg), _ptr(out, C.c_double), n.value, C.byref(n)))
        return rg.value, out

    def snapshot(self, sid):
        """The system's whole device state as bytes (exact checkpoint)."""
        n = C.c_uint64()
        self._chk(lib().fmd_snapshot_size(self.h, sid, C.byref(n)))
        buf = np.zeros(n.value, np.uint8)
        self._chk(lib().fmd_snapshot(self.h, sid, buf.ctypes.data, n.value))
        return buf.tobytes()

    def resume(self, snap):
        """A new system continuing exactly from a snapshot."""
        buf = np.frombuffer(snap, np.uint8)
        sid = C.c_int32()
        self._chk(lib().fmd_system_resume(self.h, buf.ctypes.data, len(buf), C.byref(sid)))
        self.natoms[sid.value] = lib().fmd_system_natoms(self.h, sid.value)
        return sid.value

    def corr_push(self, sid, sample6):
        """Push one stress sample (xx, xy, xz, yy, yz, zz) into the system's correlator."""
        x = np.ascontiguousarray(sample6, np.float64).reshape(6)
        self._chk(lib().fmd_corr_push(self.h, sid, _ptr(x, C.c_double)))

    def corr_reset(self, sid):
        self._chk(lib().fmd_corr_reset(self.h, sid))

    def corr(self, sid):
        """Unnormalised correlator sums (levels, points, 6) and counts (levels, points)."""
        c = np.zeros((CORR_LEVELS, CORR_POINTS, 6), np.float64)
        n = np.zeros((CORR_LEVELS, CORR_POINTS), np.int32)
        lv = C.c_int32()
        self._chk(lib().fmd_get_corr(self.h, sid, _ptr(c, C.c_double), _ptr(n, C.c_int32), C.byref(lv)))
        return c[:lv.value], n[:lv.value]

    def state_box(self, sid):
        """Box side of the positions state() returns (differs from box() under NPT between steps)."""
        L = C.c_double()
        self._chk(lib().fmd_get_state_box(self.h, sid, C.byref(L)))
        return L.value

    def thermo(self, sid):
        t = Thermo()
        self._chk(lib().fmd_get_thermo(self.h, sid, C.byref(t)))
        return t.as_dict()

    def thermostat(self, sid):
        """Nose-Hoover chain and barostat state (FP64) after the last step."""
        t = Thermostat()
        self._chk(lib().fmd_get_thermostat(self.h, sid, C.byref(t)))
        return t.as_dict()

    def nh_chain_run(self, M, Q1, Qr, xi, vxi, ke, kT, natoms, half_dt, carry=False):
        """The device chain applied to a stream of kinetic energies, state
        carried from draw to draw; returns (scales, xi, vxi)."""
        ke = np.ascontiguousarray(ke, np.float64).reshape(-1)
        xi = np.ascontiguousarray(np.resize(np.asarray(xi, np.float64), MAX_CHAIN))
        vxi = np.ascontiguousarray(np.resize(np.asarray(vxi, np.float64), MAX_CHAIN))
        sc = np.zeros(len(ke) * (2 if carry else 1), np.float64)
        c = ChainRun()
        c.M, c.natoms, c.carry, c.n = M, natoms, 1 if carry else 0, len(ke)
        c.Q1, c.Qr, c.kT, c.half_dt = Q1, Qr, kT, half_dt
        c.ke = _ptr(ke, C.c_double)
        self._chk(lib().fmd_nh_chain_run(self.h, C.byref(c), _ptr(xi, C.c_double), _ptr(vxi, C.c_double),
                                         _ptr(sc, C.c_double)))
        return sc, xi, vxi

    def drain_thermo(self, sid, cap=4096):
        buf = (Thermo * cap)()
        n = C.c_int32()
        lost = C.c_int64()
        self._chk(lib().fmd_drain_thermo(self.h, sid, buf, cap, C.byref(n), C.byref(lost)))
        return [buf[k].as_dict() for k in range(n.value)], lost.value

    def autotune(self, sids, reps=5):
        """Tunes per size class; returns {(kernel, class): value} (kernel 0 =
        force block, 1 = list-build block, 2 = neighbor layout)."""
        a = np.asarray(sids, np.int32)
        self._chk(lib().fmd_autotune(self.h, _ptr(a, C.c_int32), len(a), reps))
        return self.tuned()

    def tuned(self):
        return {(k, c): lib().fmd_tuned_block(self.h, k, c) for k in (0, 1, 2) for c in (0, 1, 2)}

    def set_tuned(self, kernel, cls, value):
        self._chk(lib().fmd_engine_set_tuned(self.h, kernel, cls, value))

    def autotune_times(self):
        """{class: [(layout, block, ms), ...]} of the last autotune."""
        out = {}
        for c in (0, 1, 2):
            n = lib().fmd_autotune_times(self.h, c, None, None, None, 0)
            if n <= 0:
                continue
            lay, blk, ms = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float64)
            lib().fmd_autotune_times(self.h, c, _ptr(lay, C.c_int32), _ptr(blk, C.c_int32), _ptr(ms, C.c_double), n)
            out[c] = [(int(a), int(b), float(m)) for a, b, m in zip(lay, blk, ms)]
        return out

    def time_force(self, sids, terms=TERM_LJ, reps=20):
        a = np.asarray(sids, np.int32)
        ms, b, p = C.c_double(), C.c_double(), C.c_int64()
        self._chk(lib().fmd_time_force(self.h, _ptr(a, C.c_int32), len(a), terms, reps, C.byref(ms),
                                       C.byref(b), C.byref(p)))
        return ms.value, b.value, p.value

    def bytes_in_use(self):
        return lib().fmd_engine_bytes_in_use(self.h)


def system_bytes(desc):
    c = desc.c if hasattr(desc, "c") else desc
    return lib().fmd_system_bytes(C.byref(c))


def size_class(natoms):
    return lib().fmd_size_class(natoms)


def admission_replay(t_arrive, nbytes, duration, budget, max_jobs=0):
    t_arrive = np.ascontiguousarray(t_arrive, np.float64)
    nbytes = np.ascontiguousarray(nbytes, np.uint64)
    duration = np.ascontiguousarray(duration, np.float64)
    n = len(t_arrive)
    ta = np.zeros(n, np.float64)
    tf = np.zeros(n, np.float64)
    rc = lib().fmd_admission_replay(n, _ptr(t_arrive, C.c_double), _ptr(nbytes, C.c_uint64),
                                    _ptr(duration, C.c_double), budget, max_jobs,
                                    _ptr(ta, C.c_double), _ptr(tf, C.c_double))
    if rc:
        raise FmdError(rc, "admission replay")
    return ta, tf


class Server:
    """Continuous-batching job server (owns its own engine on a worker thread)."""

    def __init__(self, device=0, hbm_budget=0, max_jobs=0, graph_steps=0):
        o = ServerOpts(hbm_budget, max_jobs, graph_steps)
        h = C.c_void_p()
        rc = lib().fmd_server_create(device, C.byref(o), C.byref(h))
        if rc:
            raise FmdError(rc, _host_err())
        self.h = h
        self._descs = {}

    def submit(self, desc, nsteps):
        c = desc.c if hasattr(desc, "c") else desc
        jid = C.c_int64()
        rc = lib().fmd_server_submit(self.h, C.byref(c), nsteps, C.byref(jid))
        if rc:
            raise FmdError(rc, "submit")
        self._descs[jid.value] = c.natoms
        return jid.value

    def poll(self, jid):
        s = C.c_int32()
        lib().fmd_server_poll(self.h, jid, C.byref(s))
        return s.value

    def result(self, jid, state=False):
        r = JobResult()
        n = self._descs[jid]
        p = np.zeros((n, 4), np.float32) if state else None
        v = np.zeros((n, 3), np.float32) if state else None
        rc = lib().fmd_server_result(self.h, jid, C.byref(r), _ptr(p, C.c_float), _ptr(v, C.c_float))
        if rc:
            raise FmdError(rc, "result")
        out = {"status": r.status, "natoms": r.natoms, "steps": r.steps, "t_submit": r.t_submit,
               "t_admit": r.t_admit, "t_done": r.t_done, "last": r.last.as_dict()}
        if state:
            out["pos4"], out["vel"] = p, v
        return out

    def set_blocks(self, tuned):
        """tuned: {(kernel, size_class): block}, e.g. from Engine.autotune."""
        for (kern, cls), b in tuned.items():
            rc = lib().fmd_server_set_block(self.h, kern, cls, b)
            if rc:
                raise FmdError(rc, f"server block ({kern}, {cls}) = {b}")

    def drain(self):
        lib().fmd_server_drain(self.h)

    def now(self):
        return lib().fmd_server_now(self.h)

    def stats(self):
        bs = C.c_double()
        gl = C.c_int64()
        res = C.c_int32()
        q = C.c_int32()
        lib().fmd_server_stats(self.h, C.byref(bs), C.byref(gl), C.byref(res), C.byref(q))
        return {"bead_steps": bs.value, "graph_launches": gl.value, "resident": res.value, "queued": q.value}

    def close(self):
        if getattr(self, "h", None):
            lib().fmd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
