import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def golden():
    return dict(np.load(os.path.join(GOLDEN, "ref_inputs.npz")))


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def fmd():
    import fastmd_amd
    return fastmd_amd


@pytest.fixture(scope="session")
def engine(fmd):
    e = fmd.Engine(0)
    yield e
    e.close()


def type_bits(t):
    return np.array(t, dtype=np.int32).view(np.float32)


def pos4_of(xyz, types=0):
    xyz = np.asarray(xyz, np.float32)
    out = np.zeros((len(xyz), 4), np.float32)
    out[:, :3] = xyz
    out[:, 3] = type_bits(np.broadcast_to(np.asarray(types, np.int32), (len(xyz),)))
    return out


KG_LJ = np.array([1.0, 1.0], np.float32)
KG_BP = np.array([[30.0, 1.5, 1.0, 1.0]], np.float32)
KG_AP = np.array([[5.0, np.float32(180.0) * np.float32(3.14159265358979323846) / np.float32(180.0)]], np.float32)


def relax_kg(fmd, engine, nchains, chain_len, seed, rho=0.85, **kw):
    """Random-walk melt relaxed on the GPU by the soft push-off of
    fastmd_amd.workloads, returned as a Desc of the relaxed state."""
    from fastmd_amd import workloads
    spec = {"nchains": nchains, "chain_len": chain_len, "T": 1.0, "seed": 42, "gen_seed": seed}
    sid = workloads.relaxed_replicas(engine, [spec])[0]
    p, v, im = engine.state(sid)
    engine.remove(sid)
    t = fmd.Topo.kg_melt(nchains, chain_len, rho, seed)
    a = t.arrays()
    return fmd.Desc(p, t.c.L, a["lj"], vel=v, image=im, mol=a["mol"], bonds=a["bonds"],
                    bond_types=a["bond_types"], bond_params=a["bond_params"], angles=a["angles"],
                    angle_params=a["angle_params"], masses=a["masses"], **kw)


def rel_err_floor(got, ref, floor):
    """max over atoms of |got - ref| / max(|ref|, floor) on 3-vectors."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    num = np.linalg.norm(got - ref, axis=-1)
    den = np.maximum(np.linalg.norm(ref, axis=-1), floor)
    return float(np.max(num / den))


def lattice_chains(nside, rho=0.85, seed=1, T=1.0):
    """A dense bead-spring system with no overlaps, built without a push-off:
    nside^3 beads on a simple cubic lattice at density rho (spacing about
    1.06 sigma at 0.85), each lattice row along x one straight chain of nside
    beads (bonds between neighbours in the row, angles of 180 degrees), and
    Maxwell velocities at T with zero total momentum.  Returns
    (pos4, vel, L, bonds, angles)."""
    n = nside ** 3
    L = np.float32((n / rho) ** (1.0 / 3.0))
    a = float(L) / nside
    g = np.arange(n)
    xyz = np.stack([g % nside, (g // nside) % nside, g // nside ** 2], 1) * a + 0.5 * a
    bonds = np.array([[i, i + 1] for i in range(n) if (i % nside) != nside - 1], np.int32)
    angles = np.array([[i, i + 1, i + 2, 0] for i in range(n) if (i % nside) < nside - 2], np.int32)
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, np.sqrt(T), (n, 3))
    v -= v.mean(0)
    return pos4_of(xyz), v.astype(np.float32), L, bonds, angles
