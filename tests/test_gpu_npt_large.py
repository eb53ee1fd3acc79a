"""Mixed-precision NPT at scale: a 5,000 x 100-bead Kremer-Grest melt
(500,000 beads) under Nose-Hoover NPT (chain 3, Tdamp 0.1, P ramp 0 -> 1,
Pdamp 1).  Pair, bond and angle forces are FP32 on the device; kinetic and
virial reductions and the thermostat/barostat are FP64.  After NPT steps the
forces, potential energy and pressure tensor are compared with the FP64
oracle on the same state, each quantity's error stated and bounded at 1e-5
relative (src/integrate/nose_hoover.cu:606-757 for the operator order)."""
import numpy as np
import pytest

from conftest import KG_AP, KG_BP, KG_LJ, rel_err_floor, relax_kg

pytestmark = pytest.mark.gpu

TOL = 1e-5


def test_npt_500k_error_per_quantity(fmd, engine, orc, capsys):
    d = relax_kg(fmd, engine, 5000, 100, seed=21, ensemble=fmd.NPT_NH, T=(1.0, 1.0), Tdamp=0.1,
                 P=(0.0, 1.0), Pdamp=1.0, chain=3, nsteps=20000, thermo_freq=50)
    sid = engine.add(d)
    engine.step(sid, 200)
    L0 = float(d.c.L)
    L = np.float32(engine.state_box(sid))  # the box the returned positions are in
    assert L != np.float32(L0)  # the barostat moved the box
    p4, v, _ = engine.state(sid)
    f, vir, pe = engine.forces(sid)
    f8, v8 = orc.forces_f64(p4, L, KG_LJ, bonds=d.bonds, btypes=d.bond_types, bp=KG_BP, ang4=d.angles,
                            ap=KG_AP, terms=7)
    floor = np.sqrt(np.mean(np.sum(f8[:, :3] ** 2, axis=1)))
    err_f = rel_err_floor(f[:, :3], f8[:, :3], floor)
    pe8 = f8[:, 3].sum()
    err_pe = abs(pe - pe8) / abs(pe8)
    err_w = np.max(np.abs(vir - v8)) / np.abs(v8).max()
    # pressure tensor P_ab = (sum m v_a v_b + W_ab) / V with unit masses
    vd = v.astype(np.float64)
    kin = np.array([vd[:, 0] @ vd[:, 0], vd[:, 0] @ vd[:, 1], vd[:, 0] @ vd[:, 2],
                    vd[:, 1] @ vd[:, 1], vd[:, 1] @ vd[:, 2], vd[:, 2] @ vd[:, 2]])
    V = float(L) ** 3
    p_dev, p_ref = (kin + vir) / V, (kin + v8) / V
    # P is a near-cancelling sum (|P| << kinetic and virial parts near the P = 0
    # start of the ramp), so its error is stated relative to the parts' scale
    scale_p = np.max(np.abs(kin[[0, 3, 5]]) + np.abs(v8[[0, 3, 5]])) / V
    err_p = np.max(np.abs(p_dev - p_ref)) / scale_p
    with capsys.disabled():
        print(f"\n[npt 500k] L {L0:.4f} -> {float(L):.4f}; error vs FP64 oracle: force {err_f:.2e}, "
              f"PE {err_pe:.2e}, virial {err_w:.2e}, pressure tensor {err_p:.2e} "
              f"(abs {np.max(np.abs(p_dev - p_ref)):.2e}, P_xx {p_ref[0]:.4f})")
    assert err_f < TOL and err_pe < TOL and err_w < TOL and err_p < TOL
    th = engine.thermo(sid)
    assert np.isfinite(th["temp"]) and 0.7 < th["temp"] < 1.3
    engine.remove(sid)
