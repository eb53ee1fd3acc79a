"""Tabulated pair potentials on the device (src/force/table.cu:6-182) over the
same Verlet rows as LJ: parity with the FP64 oracle's linear interpolant at
1e-5 relative, the reference's LJ-table check (tests/test_table.cu:8-115), a
mixed LJ + table melt, and dynamics with a table."""
import numpy as np
import pytest

from conftest import KG_LJ, pos4_of, rel_err_floor, relax_kg
from test_oracle_pins import lj_table_rows

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _check(f, vir, pe, f8, v8, tol=TOL):
    floor = np.sqrt(np.mean(np.sum(f8[:, :3] ** 2, axis=1)))
    err_f = rel_err_floor(f[:, :3], f8[:, :3], floor)
    pe8 = f8[:, 3].sum()
    err_pe = abs(pe - pe8) / max(abs(pe8), 1e-3 * len(f8))
    err_w = np.max(np.abs(vir - v8)) / np.abs(v8).max()
    assert err_f < tol and err_pe < tol and err_w < tol, (err_f, err_pe, err_w)
    return err_f, err_pe, err_w


def test_lj_table_on_reference_input(fmd, engine, orc, golden):
    # tests/test_table.cu:8-115: 64 random atoms in L = 8, a 10,000-point LJ
    # table with the force/r clamp against the LJ kernel, 1e-3 on forces
    p4 = pos4_of(golden["lj64_xyz"])
    L = np.float32(8.0)
    rows = lj_table_rows()
    tabs = ([0], [(0.01, 2.5, rows)])
    sid_lj = engine.add(fmd.Desc(p4, L, KG_LJ, rc=2.5, skin=0.5))
    sid_tb = engine.add(fmd.Desc(p4, L, np.zeros(2, np.float32), rc=2.5, skin=0.5, tables=tabs))
    fl, vl, pl = engine.forces(sid_lj, fmd.TERM_LJ)
    ft, vt, pt = engine.forces(sid_tb, fmd.TERM_LJ)
    print(f"[table] reference input: max |F_table - F_lj| {np.abs(ft[:, :3] - fl[:, :3]).max():.2e}", flush=True)
    assert np.abs(ft[:, :3] - fl[:, :3]).max() < 1e-3
    rel_pe = np.abs(ft[:, 3] - fl[:, 3]) / np.maximum(np.abs(fl[:, 3]), 1e-6)
    assert rel_pe.max() < 1e-3
    assert np.max(np.abs(vt - vl)) < 1e-2 * np.abs(vl).max()
    f8, v8 = orc.forces_f64(p4, L, np.zeros(2, np.float32), terms=1, tables=tabs)
    _check(ft, vt, pt, f8, v8)
    engine.remove(sid_lj)
    engine.remove(sid_tb)


def _mixed_desc(fmd, d, types, table_rows, **kw):
    """The relaxed melt d with per-chain types: LJ on (0, 0) and (1, 1), the
    table on (0, 1)."""
    p4 = d.pos4.copy()
    p4[:, 3] = np.asarray(types, np.int32).view(np.float32)
    lj = np.array([[1.0, 1.0], [0.0, 0.0], [0.0, 0.0], [1.0, 1.0]], np.float32)
    tabs = ([-1, 0, 0, -1], [(0.5, 2.5, table_rows)])
    desc = fmd.Desc(p4, d.c.L, lj, ntypes=2, vel=d.vel, image=d.image, mol=d.mol, bonds=d.bonds,
                    bond_types=d.bond_types, bond_params=d.bond_params, angles=d.angles,
                    angle_params=d.angle_params, masses=np.ones(2, np.float32), tables=tabs, **kw)
    return desc, lj, tabs


def _soft_rows(n=2000):
    # a repulsive-attractive tabulated pair that is not LJ: e = 3 exp(-2 (r - 0.8)) - 1 / r^4
    r = np.linspace(0.5, 2.5, n)
    e = 3.0 * np.exp(-2.0 * (r - 0.8)) - r ** -4
    fr = (6.0 * np.exp(-2.0 * (r - 0.8)) - 4.0 * r ** -5) / r  # -dE/dr / r
    return np.stack([r, e, fr], axis=1).astype(np.float32)


def test_mixed_lj_and_table_melt_matches_oracle(fmd, engine, orc, capsys):
    d = relax_kg(fmd, engine, 60, 50, seed=23)
    types = np.repeat(np.arange(60) % 2, 50)
    desc, lj, tabs = _mixed_desc(fmd, d, types, _soft_rows())
    sid = engine.add(desc)
    for k in range(2):
        if k:
            engine.step(sid, 100)
        p4, v, _ = engine.state(sid)
        L = np.float32(engine.state_box(sid))
        f, vir, pe = engine.forces(sid, fmd.TERM_ALL)
        f8, v8 = orc.forces_f64(p4, L, lj, ntypes=2, bonds=d.bonds, btypes=d.bond_types, bp=d.bond_params,
                                ang4=d.angles, ap=d.angle_params, terms=7, tables=tabs)
        errs = _check(f, vir, pe, f8, v8)
        with capsys.disabled():
            print(f"\n[table] mixed melt after {100 * k} steps: force {errs[0]:.2e}, PE {errs[1]:.2e}, "
                  f"virial {errs[2]:.2e}", flush=True)
    th = engine.thermo(sid)
    assert np.isfinite(th["pe"]) and 0.7 < th["temp"] < 1.3
    engine.remove(sid)


def test_table_only_pairs_in_16_and_32_bit_rows(fmd, engine, orc, monkeypatch):
    # a table on every pair (one type); FMD_IDX32 selects 32-bit neighbor rows,
    # which must give the same forces bit for bit
    d = relax_kg(fmd, engine, 40, 50, seed=29)
    tabs = ([0], [(0.01, 2.5, lj_table_rows())])
    desc = fmd.Desc(d.pos4, d.c.L, np.zeros(2, np.float32), vel=d.vel, image=d.image, mol=d.mol, bonds=d.bonds,
                    bond_types=d.bond_types, bond_params=d.bond_params, angles=d.angles,
                    angle_params=d.angle_params, tables=tabs)
    sids = {}
    for wide in (False, True):
        if wide:
            monkeypatch.setenv("FMD_IDX32", "1")
        else:
            monkeypatch.delenv("FMD_IDX32", raising=False)
        sids[wide] = engine.add(desc)
    monkeypatch.delenv("FMD_IDX32", raising=False)
    p4, _, _ = engine.state(sids[False])
    L = np.float32(engine.state_box(sids[False]))
    f8, v8 = orc.forces_f64(p4, L, np.zeros(2, np.float32), bonds=d.bonds, btypes=d.bond_types,
                            bp=d.bond_params, ang4=d.angles, ap=d.angle_params, terms=7, tables=tabs)
    f, vir, pe = engine.forces(sids[False], fmd.TERM_ALL)
    _check(f, vir, pe, f8, v8)
    for a, b in zip(engine.forces(sids[False]), engine.forces(sids[True])):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    for s in sids.values():
        engine.remove(s)
