"""Kremer-Grest melt workloads: generation on the host, soft push-off on the GPU.

The push-off caps the LJ pair force and ramps the cap (5 -> 1000) under strong
Langevin damping (Tdamp = 0.1), removes the cap while still strongly damped (so
the energy of the last overlaps is drained), then runs at the production
temperature and damping.  All replicas of a workload are relaxed together in
batched steps.
"""
import numpy as np

from .api import LANGEVIN, Desc, Topo

PUSH_OFF = ((5.0, 200), (20.0, 200), (100.0, 200), (1000.0, 300))
COOL = 400
FINAL = 200


def uneven_replica_specs(n=256, seed=12345, beads=(2000, 64000), chain=(20, 200), temp=(0.8, 1.2),
                         lang_seed0=1000, gen_seed0=2000):
    """Uneven replicas: chain length and total beads log-uniform, T uniform."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n):
        clen = int(round(np.exp(rng.uniform(np.log(chain[0]), np.log(chain[1])))))
        b = float(np.exp(rng.uniform(np.log(beads[0]), np.log(beads[1]))))
        nch = max(1, int(round(b / clen)))
        T = float(rng.uniform(temp[0], temp[1]))
        out.append({"nchains": nch, "chain_len": clen, "T": T, "seed": lang_seed0 + r,
                    "gen_seed": gen_seed0 + r})
    return out


def melt_desc(spec, rho=0.85, ensemble=LANGEVIN, nsteps=1000, thermo_freq=0, force_cap=0.0, Tdamp=1.0,
              **kw):
    t = Topo.kg_melt(spec["nchains"], spec["chain_len"], rho, spec["gen_seed"])
    a = t.arrays()
    T = spec.get("T", 1.0)
    return Desc(a["pos4"], t.c.L, a["lj"], vel=a["vel"], image=a["image"], mol=a["mol"], bonds=a["bonds"],
                bond_types=a["bond_types"], bond_params=a["bond_params"], angles=a["angles"],
                angle_params=a["angle_params"], masses=a["masses"], ensemble=ensemble, T=(T, T), Tdamp=Tdamp,
                seed=spec.get("seed", 42), nsteps=nsteps, thermo_freq=thermo_freq, force_cap=force_cap, **kw)


def state_desc(topo_desc, pos4, vel, image, L, ensemble=LANGEVIN, T=1.0, Tdamp=1.0, seed=42, nsteps=1000):
    """A descriptor with the topology of `topo_desc` (a Desc) and the given
    state, ensemble, temperature and seed: how a relaxed melt is resubmitted
    as a new simulation job."""
    t = topo_desc
    return Desc(pos4, L, t.lj, vel=vel, image=image, mol=t.mol, bonds=t.bonds, bond_types=t.bond_types,
                bond_params=t.bond_params, angles=t.angles, angle_params=t.angle_params, masses=t.masses,
                ensemble=ensemble, T=(T, T), Tdamp=Tdamp, seed=seed, nsteps=nsteps)


def job_stream(pool_natoms, rate_bead_steps, window_s, steps=(1000, 100000), load=0.8, nvt_frac=0.3,
               temp=(0.8, 1.2), seed=777):
    """Poisson job stream over a pool of starting melts.

    Each job draws its melt uniformly from the pool, its run length
    log-uniformly in `steps`, its ensemble (Nose-Hoover NVT with probability
    `nvt_frac`, else Langevin) and T uniformly in `temp`.  The arrival rate is
    fixed at load * rate_bead_steps / E[bead-steps per job].  Returns
    (rate in jobs/s, list of job dicts with arrival time t)."""
    rng = np.random.default_rng(seed)
    a, b = steps
    e_steps = (b - a) / np.log(b / a)
    e_bs = float(np.mean(pool_natoms)) * e_steps
    lam = load * rate_bead_steps / e_bs
    jobs, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / lam)
        if t >= window_s:
            break
        jobs.append({"t": t, "pool": int(rng.integers(len(pool_natoms))),
                     "steps": int(round(np.exp(rng.uniform(np.log(a), np.log(b))))),
                     "nvt": bool(rng.random() < nvt_frac), "T": float(rng.uniform(*temp)),
                     "seed": int(rng.integers(1, 2**31 - 1))})
    return lam, jobs


def relaxed_replicas(engine, specs, stages=PUSH_OFF, cool=COOL, final=FINAL, Tdamp=1.0, log=None,
                     descs=None, rho=0.85):
    """Creates the replicas (melts at number density rho) on the engine,
    relaxes them in batched steps and returns their system ids (Langevin at
    each replica's T, Tdamp = 1).  If `descs` is a list, the generated
    descriptors are appended to it."""
    sids = []
    for s in specs:
        d = melt_desc(s, rho=rho, force_cap=stages[0][0], Tdamp=0.1)
        sids.append(engine.add(d))
        if descs is not None:
            descs.append(d)
    for cap, n in stages:
        for sid in sids:
            engine.set_force_cap(sid, cap)
        engine.step(sids, n)
        if log:
            log(f"push-off cap {cap}: {n} steps")
    for sid, s in zip(sids, specs):
        engine.set_force_cap(sid, 0.0)
        engine.set_langevin(sid, s["T"], s["T"], 0.1)
    if cool:
        engine.step(sids, cool)
    for sid, s in zip(sids, specs):
        engine.set_langevin(sid, s["T"], s["T"], Tdamp)
    if final:
        engine.step(sids, final)
    return sids
