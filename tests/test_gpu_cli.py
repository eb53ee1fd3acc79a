"""End-to-end drop-in: the fastmd binary run on a run.conf in the reference's
key format (tests/integration/bonded_ring/fastmd_template.conf), writing the
reference's thermo, Rg, trajectory, stress-correlation and restart files and stdout lines
(src/main.cu:403-811; benchmarks/run.py:44-57 parses the stdout)."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, relax_kg

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "fastmd_amd", "bin", "fastmd")


def test_cli_run_conf_outputs(fmd, engine, tmp_path):
    d = relax_kg(fmd, engine, 40, 50, seed=41)
    sid = engine.add(d)
    p4, v, im = engine.state(sid)
    L = engine.box(sid)
    engine.remove(sid)
    t = fmd.Topo.kg_melt(40, 50, 0.85, 41)
    data = tmp_path / "melt.data"
    t.write_data(str(data), pos4=p4, vel=v, image=im, L=L, step=0)
    conf = f"""natoms 2000
ntypes 1
rc 2.5
skin 0.3
dt 0.001
nvt_langevin 1.0 1.0 1.0 42
nsteps 400
dump_freq 100
thermo 1 50 thermo.dat
rg 1 100 rg.dat
stress 1 10 stress.dat
lj 0 0 1.0 1.0
bond_type 0 30.0 1.5 1.0 1.0
angle_type 0 5.0 180.0
lammps_data_file {data}
restart 200 restart.data
"""
    (tmp_path / "run.conf").write_text(conf)
    print(f"[cli] running {BIN} run.conf in {tmp_path}", flush=True)
    r = subprocess.run([BIN, "run.conf"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = r.stdout
    # stdout contract parsed by benchmarks/run.py
    assert re.search(r"Wall time: [0-9.]+ s\s+\|\s+Steps/s: [0-9.]+\s+\|\s+ns/day: [0-9.]+", out)
    assert len(re.findall(r"^Step\s+\d+\s+T=", out, re.M)) == 8
    # maxV / maxF of the thermo line (src/main.cu:645-648): a relaxed melt at
    # T = 1 has its fastest bead at a few sigma/tau and its largest force well
    # above zero; the spike lines appear only above 10 and 100
    mv = [float(x) for x in re.findall(r"maxV=\s*([0-9.]+)", out)]
    mf = [float(x) for x in re.findall(r"maxF=\s*([0-9.]+)", out)]
    assert len(mv) == 8 and len(mf) == 8
    assert all(2.0 < x < 10.0 for x in mv) and all(10.0 < x for x in mf)
    assert ("FORCE SPIKE" in out) == any(x > 100.0 for x in mf) and "VELOCITY SPIKE" not in out
    # thermo file: header and 10 columns, every 50 steps
    rows = [l for l in (tmp_path / "thermo.dat").read_text().splitlines()]
    assert rows[0] == "# step  KE  PE  T  Pxx  Pyy  Pzz  Pxy  Pxz  Pyz"
    vals = np.array([list(map(float, l.split())) for l in rows[1:]])
    assert vals.shape == (8, 10) and list(vals[:, 0]) == [50 * k for k in range(1, 9)]
    assert np.all((vals[:, 3] > 0.7) & (vals[:, 3] < 1.3))
    # Rg file every 100 steps
    rg = (tmp_path / "rg.dat").read_text().splitlines()
    assert rg[0] == "# step  Rg" and len(rg) == 5
    assert [int(l.split()[0]) for l in rg[1:]] == [100, 200, 300, 400]
    assert all(2.0 < float(l.split()[1]) < 6.0 for l in rg[1:])
    # trajectory frames every 100 steps (src/io/dump.cu layout)
    raw = (tmp_path / "traj.bin").read_bytes()
    hdr = np.frombuffer(raw[:12], np.int32)
    assert list(hdr) == [0x4D444247, 2000, 1]
    fsz = 8 + 4 + 4 + 12 * 2000
    assert len(raw) == 12 + 4 * fsz
    for k in range(4):
        off = 12 + k * fsz
        assert np.frombuffer(raw[off:off + 8], np.int64)[0] == 100 * (k + 1)
        assert np.frombuffer(raw[off + 8:off + 12], np.int32)[0] == 2000
        Lk = np.frombuffer(raw[off + 12:off + 16], np.float32)[0]
        xyz = np.frombuffer(raw[off + 16:off + fsz], np.float32).reshape(2000, 3)
        assert abs(Lk - L) < 1e-4 and xyz.min() >= 0.0 and xyz.max() <= Lk
    # stress autocorrelation (src/main.cu:758-783): 40 samples every 10 steps;
    # level 0 has 16 lags, level 1 (means of 16) two samples -> two lags
    assert "Stress autocorrelation written to stress.dat" in out
    st = (tmp_path / "stress.dat").read_text().splitlines()
    assert st[0] == "# lag_steps  C_xx  C_xy  C_xz  C_yy  C_yz  C_zz  count"
    cols = [l.split() for l in st[1:]]
    assert all(len(c) == 8 for c in cols)
    assert [int(c[0]) for c in cols] == [10 * (k + 1) for k in range(16)] + [160, 320]
    lag = np.arange(16)
    assert [int(c[7]) for c in cols] == list(2 * (16 - lag) + np.maximum(0, 8 - lag)) + [2, 1]
    assert all(float(c[1]) > 0 for c in cols[:1])  # <P_xx^2> at lag 0
    # restart files read back with the same topology
    for name in ("restart_200.data", "restart_final.data"):
        t2 = fmd.Topo.from_data(str(tmp_path / name))
        assert t2.natoms == 2000 and len(t2.arrays()["bonds"]) == 40 * 49


def test_cli_table_potential_matches_engine(fmd, engine, orc, tmp_path):
    """run.conf with a `table` pair (src/io/config.cpp:136-162) and a
    forces_dump_file: the CLI's initial forces dump (src/main.cu:344-362)
    equals the FP64 oracle with the same table."""
    from test_oracle_pins import lj_table_rows
    d = relax_kg(fmd, engine, 10, 40, seed=43)
    sid = engine.add(d)
    p4, v, im = engine.state(sid)
    L = engine.box(sid)
    engine.remove(sid)
    t = fmd.Topo.kg_melt(10, 40, 0.85, 43)
    data = tmp_path / "melt.data"
    t.write_data(str(data), pos4=p4, vel=v, image=im, L=L, step=0)
    rows = lj_table_rows(n=4000, rlo=0.5)
    with open(tmp_path / "lj.table", "w") as f:
        f.write("# LJ 12-6 with the force/r clamp\nLJ_TAB\nN %d R %.9g %.9g\n\n" % (len(rows), rows[0, 0], rows[-1, 0]))
        for k, (r, e, fr) in enumerate(rows):
            f.write("%d %.9g %.9g %.9g\n" % (k + 1, r, e, fr))
    conf = f"""natoms 400
ntypes 1
rc 2.5
skin 0.3
dt 0.001
nvt_langevin 1.0 1.0 1.0 42
nsteps 50
table 0 0 {tmp_path / "lj.table"} LJ_TAB
bond_type 0 30.0 1.5 1.0 1.0
angle_type 0 5.0 180.0
lammps_data_file {data}
forces_dump_file forces.txt
"""
    (tmp_path / "run.conf").write_text(conf)
    r = subprocess.run([BIN, "run.conf"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Initial forces written to forces.txt" in r.stdout
    dump = np.loadtxt(tmp_path / "forces.txt", comments="#")
    # the system as the CLI parsed it (positions from the data file, global-id order)
    tc = fmd.Topo.from_conf(str(tmp_path / "run.conf"))
    a = tc.arrays()
    assert a["table_n"].tolist() == [len(rows)] and np.array_equal(a["table_data"], rows)
    rng = a["table_range"][0]
    f8, _ = orc.forces_f64(a["pos4"], np.float32(tc.c.L), np.zeros(2, np.float32), bonds=a["bonds"],
                           btypes=a["bond_types"], bp=a["bond_params"], ang4=a["angles"], ap=a["angle_params"],
                           terms=7, tables=([0], [(float(rng[0]), float(rng[1]), a["table_data"])]))
    ids = dump[:, 0].astype(int)
    floor = np.sqrt(np.mean(np.sum(f8[:, :3] ** 2, axis=1)))
    got = dump[:, 1:4]
    ref = f8[ids, :3]
    err = np.max(np.linalg.norm(got - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), floor))
    assert err < 1e-5, err
