/* fmd.h -- C ABI of the MI355X-native bead-spring MD engine (libfastmd.so).
 *
 * The boundary the fastMD binary's step loop crosses, re-expressed as plain C:
 * opaque handles, host pointers and sizes, int32 status codes (0 = ok, never
 * exit()).  Every entry point names the reference interface it replaces.
 * The engine owns all device memory; callers pass and receive host arrays.
 *
 * Layout conventions kept from the reference (src/core/types.cuh:23-41):
 *   positions are float4 (x, y, z, w) with the int32 type id bit-cast into w;
 *   per-atom force output is float4 (fx, fy, fz, pe);
 *   virial / pressure tensors are ordered xx, xy, xz, yy, yz, zz.
 * Arrays exchanged with the host are indexed by the atom's index in the input
 * (its "global id"), never by the engine's internal cell-sorted slot.
 */
#ifndef FMD_FMD_H
#define FMD_FMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference aborts through CUDA_CHECK
 * (src/core/types.cuh:7-15) or throws std::runtime_error
 * (src/io/config.cpp:23,37-44); here every failure is a return value. */
enum {
    FMD_OK = 0,
    FMD_ERR_ARG = 1,          /* invalid argument or handle                 */
    FMD_ERR_IO = 2,           /* file cannot be opened / written            */
    FMD_ERR_PARSE = 3,        /* malformed run.conf or LAMMPS data file     */
    FMD_ERR_HIP = 4,          /* HIP runtime error                          */
    FMD_ERR_NOMEM = 5,        /* device allocation failed                   */
    FMD_ERR_BUDGET = 6,       /* system does not fit the HBM budget         */
    FMD_ERR_NBR_OVERFLOW = 7, /* neighbor count above capacity (never truncated) */
    FMD_ERR_STRETCH = 8,      /* FENE bond stretched to r >= R0             */
    FMD_ERR_NAN = 9,          /* non-finite position or velocity            */
    FMD_ERR_STATE = 10,       /* call not valid in the current state        */
    FMD_ERR_CELL = 11         /* box grew beyond the cell-grid capacity     */
};

/* Ensemble (src/core/types.cuh:43). */
enum { FMD_LANGEVIN = 0, FMD_NVT_NH = 1, FMD_NPT_NH = 2 };

/* Force terms for fmd_compute_forces. */
enum { FMD_TERM_LJ = 1, FMD_TERM_BOND = 2, FMD_TERM_ANGLE = 4, FMD_TERM_ALL = 7 };

typedef struct fmd_engine fmd_engine;
typedef struct fmd_topo fmd_topo;
typedef struct fmd_server fmd_server;

/* One simulation system (replica).  Mirrors SimParams (src/core/types.cuh:45-86)
 * plus the TopologyData arrays (src/io/config.hpp:8-31).  All pointers are host
 * memory and are copied by fmd_system_create. */
typedef struct fmd_system_desc {
    int32_t natoms;
    int32_t ntypes;
    const float* pos;          /* natoms*4: x, y, z, type bits                    */
    const float* vel;          /* natoms*3 or NULL (zero)                          */
    const int32_t* image;      /* natoms*3 or NULL (zero)                          */
    const int32_t* mol_id;     /* natoms or NULL                                   */
    int32_t nbonds;
    const int32_t* bonds;      /* nbonds*2, 0-based atom indices                   */
    const int32_t* bond_types; /* nbonds, 0-based                                  */
    int32_t nangles;
    const int32_t* angles;     /* nangles*4: i, j(center), k, type                 */
    const float* masses;       /* ntypes or NULL (all 1)                           */
    const float* lj;           /* ntypes*ntypes*2: eps, sig                        */
    int32_t nbond_types;
    const float* bond_params;  /* nbond_types*4: k, R0, eps, sig (FENEParams)      */
    int32_t nangle_types;
    const float* angle_params; /* nangle_types*2: k, theta0 in radians             */
    float L;                   /* cubic box side                                   */
    float rc, skin, dt;
    int32_t ensemble;          /* FMD_LANGEVIN / FMD_NVT_NH / FMD_NPT_NH           */
    float T_start, T_stop, Tdamp;
    float P_start, P_stop, Pdamp;
    int32_t chain_length;      /* Nose-Hoover chain M (<= 10)                      */
    uint64_t seed;             /* Langevin noise seed                              */
    int64_t nsteps;            /* run length; denominator of the T/P ramps         */
    int32_t thermo_freq;       /* 0 = no thermo records                            */
    float force_cap;           /* > 0: soft push-off: cap |F_pair| of LJ and limit the */
                               /* displacement per step to 0.02 sigma; 0 = off         */
    /* Tabulated pair potentials (src/force/table.cu:6-182), added to LJ on the
     * same neighbor rows; a type pair has either an LJ entry or a table. */
    int32_t ntables;
    const int32_t* table_idx;  /* ntypes*ntypes: table of the pair, -1 = none      */
    const int32_t* table_n;    /* ntables: points per table                        */
    const float* table_range;  /* ntables*2: rmin, rmax                            */
    const float* table_data;   /* sum(table_n)*3: r, energy, force/r per point     */
    /* Multiple-tau stress autocorrelation (src/analysis/correlator.cu:33-92):
     * the pressure tensor is pushed every stress_freq steps; 0 = off. */
    int32_t stress_freq;
} fmd_system_desc;

/* Multiple-tau correlator shape (src/analysis/correlator.cuh:4-6). */
#define FMD_CORR_LEVELS 16
#define FMD_CORR_POINTS 16

/* One thermo sample (src/analysis/thermo.cu:140-201): per-atom KE and PE,
 * T = 2 KE/(3N), pressure tensor P_ab = (sum m v_a v_b + W_ab)/V, and the
 * largest per-atom |v| and |F| with that atom's global id (max_vel_force_kernel,
 * src/analysis/thermo.cu:26-75; here the smallest id on ties, -1 if none). */
typedef struct fmd_thermo {
    int64_t step;
    double ke, pe, temp;
    double press[6];           /* xx, xy, xz, yy, yz, zz */
    double L;
    double max_v, max_f;
    int64_t max_v_id, max_f_id;
} fmd_thermo;

/* ----- library ----------------------------------------------------------- */
const char* fmd_version(void);
/* Human-readable text of the last error of this thread (host-only calls). */
const char* fmd_last_host_error(void);

/* ----- host-side input: run.conf and LAMMPS data (no GPU needed) ---------
 * fmd_topo_load_conf replaces parse_config + parse_lammps_data + finalize_params
 * (src/io/config.cpp:10-220, src/io/lammps_data.cpp:18-229); same keys,
 * same atom styles, same defaults; errors are returned, not thrown. */
int32_t fmd_topo_load_conf(const char* run_conf, fmd_topo** out);
/* parse_lammps_data alone (src/io/lammps_data.cpp:18-229). */
int32_t fmd_topo_load_data(const char* data_file, fmd_topo** out);
/* Kremer-Grest melt generator: nchains linear chains of chain_len beads at
 * number density rho, random walks with bond length 0.97 from a uniform start,
 * KG force field (LJ 1/1/2.5, FENE 30/1.5, angle k=5 theta0=180 deg). */
int32_t fmd_topo_kg_melt(int32_t nchains, int32_t chain_len, double rho, uint64_t gen_seed,
                         fmd_topo** out);
/* Fills a system descriptor pointing into the topo's arrays (valid until free). */
int32_t fmd_topo_desc(fmd_topo* t, fmd_system_desc* desc);
/* run.conf extras not in the descriptor (thermo file, forces_dump_file ...). */
typedef struct fmd_conf_extra {
    int32_t thermo_on, thermo_freq;
    char thermo_file[256];
    int32_t restart_freq;
    char restart_file[256];
    char forces_dump_file[256];
    int32_t dump_freq;
    int32_t rg_on, rg_freq;
    char rg_file[256];
    int32_t stress_on, stress_freq;
    char stress_file[256];
} fmd_conf_extra;
int32_t fmd_topo_conf_extra(const fmd_topo* t, fmd_conf_extra* out);
/* 1-2 exclusions, CSR by atom, sorted unique (build_exclusions,
 * src/io/lammps_data.cpp:231-253).  off has natoms+1 entries; returns count. */
int64_t fmd_topo_exclusions(const fmd_topo* t, int32_t* off, int32_t* lst, int64_t cap);
/* Writes a LAMMPS data file ("Atoms # molecular" + Velocities/Bonds/Angles)
 * (write_lammps_data, src/io/lammps_data.cpp:269-385). */
int32_t fmd_topo_write_data(const fmd_topo* t, const char* path, const float* pos4,
                            const float* vel3, const int32_t* image3, float L, int64_t step);
void fmd_topo_free(fmd_topo* t);

/* ----- engine ------------------------------------------------------------ */
/* One engine per device and host thread; hbm_budget bounds the memory all its
 * systems may hold (0 = 90% of free HBM). */
int32_t fmd_engine_create(int32_t device, uint64_t hbm_budget, fmd_engine** out);
void fmd_engine_destroy(fmd_engine* e);
const char* fmd_last_error(const fmd_engine* e);
/* Estimated device bytes of a system, the admission unit of the HBM budget. */
uint64_t fmd_system_bytes(const fmd_system_desc* d);
uint64_t fmd_engine_bytes_in_use(const fmd_engine* e);

/* Creates a device-resident system: replaces System::allocate + uploads +
 * VerletList/MortonSorter/LangevinState/NoseHoover init (src/main.cu:61-165)
 * and the initial build + force evaluation (src/main.cu:311-342). */
int32_t fmd_system_create(fmd_engine* e, const fmd_system_desc* d, int32_t* sys);
int32_t fmd_system_destroy(fmd_engine* e, int32_t sys);

/* Parameter changes between runs (soft push-off protocol of the melt
 * generator): LJ pair-force cap (0 = off) and Langevin T ramp / damping.
 * No counterpart in the reference, whose parameters are fixed per run.conf. */
int32_t fmd_system_set_force_cap(fmd_engine* e, int32_t sys, float force_cap);
int32_t fmd_system_set_langevin(fmd_engine* e, int32_t sys, float T_start, float T_stop, float Tdamp);

/* ----- kernel-level entry points on engine-owned state (parity) --------- */
/* Cell binning + Morton-ordered re-sort + Verlet list build
 * (VerletList::build, src/neighbor/verlet_list.cu:337-376;
 *  MortonSorter::sort_and_permute, src/core/morton.cu:65-85). */
int32_t fmd_build_neighbors(fmd_engine* e, int32_t sys);
/* The current Verlet pair set as ordered (i, j) global-id pairs sorted by i
 * then j (the set src/neighbor/verlet_list.cu:108-335 stores in nbr[k*N+i]).
 * Writes min(total, cap) pairs to ij (2 ints each); *npairs = total. */
int32_t fmd_get_pairs(fmd_engine* e, int32_t sys, int64_t* npairs, int32_t* ij, int64_t cap);
/* Reference cell id per atom, (cz*nx+cy)*nx+cx (assign_cells,
 * src/neighbor/verlet_list.cu:58-75), computed on the device. */
int32_t fmd_get_cell_ids(fmd_engine* e, int32_t sys, int32_t* cell_ids, int32_t* nx);
/* 30-bit Morton key per atom (compute_morton_keys, src/core/morton.cu:12-28). */
int32_t fmd_get_morton_keys(fmd_engine* e, int32_t sys, uint32_t* keys);
/* Forces, per-atom PE, virial and total PE on the current positions with the
 * current list: replaces launch_lj_kernel (src/force/lj.cuh:4-12),
 * launch_fene_kernel (src/force/fene.cuh:11-19) and launch_angle_kernel
 * (src/force/angle.cuh:9-15).  force4 may be NULL. */
int32_t fmd_compute_forces(fmd_engine* e, int32_t sys, uint32_t terms, float* force4,
                           double* virial6, double* pe);

/* ----- stepping -------------------------------------------------------- */
/* Batched step: advances every listed system by nsteps steps in hipGraph-
 * captured multi-replica launches (replaces the per-system loop of
 * src/main.cu:403-736).  Systems that reach an error stop; the first error
 * code is returned. */
int32_t fmd_step(fmd_engine* e, const int32_t* sys, int32_t nsys, int64_t nsteps);
/* The plain path: one system, one launch per kernel per step, no graph. */
int32_t fmd_step_plain(fmd_engine* e, int32_t sys, int64_t nsteps);
/* Steps completed so far. */
int64_t fmd_system_step(fmd_engine* e, int32_t sys);
/* The error that stopped a system (FMD_OK while it is healthy).  No reference
 * counterpart: the reference truncates the neighbor list
 * (src/neighbor/verlet_list.cu:281-285), clamps an overstretched FENE bond
 * (src/force/fene.cu:33) and carries NaNs on.  Here a system stops on its own,
 * the others of its batch continue unchanged:
 *   FMD_ERR_NAN: a non-finite position or velocity after the position update
 *     of step s; the system stops before that step's list build and forces
 *     (fmd_system_step = s - 1, the state as that update left it);
 *   FMD_ERR_NBR_OVERFLOW (list build of step s) and FMD_ERR_STRETCH (FENE
 *     r >= R0 in the forces of step s): the system stops at the end of step s
 *     (fmd_system_step = s).  Within a Langevin run the velocities of such a
 *     state already hold the next step's first half kick, which the force
 *     pass applies together with the forces.
 * A stopped system is not stepped again; it can be read and destroyed. */
int32_t fmd_system_error(fmd_engine* e, int32_t sys);
/* Bead count of a system (-1 if there is no such system). */
int32_t fmd_system_natoms(fmd_engine* e, int32_t sys);
/* State by global id: pos4 (x, y, z, type bits), vel3, image3; any may be NULL. */
int32_t fmd_get_state(fmd_engine* e, int32_t sys, float* pos4, float* vel3, int32_t* image3);
/* Current box side (changes under NPT): as in the reference, the barostat
 * advances it after the forces of a step (src/integrate/nose_hoover.cu:185-232),
 * and the positions follow at the next step's position update. */
int32_t fmd_get_box(fmd_engine* e, int32_t sys, double* L);
/* Box side the positions returned by fmd_get_state were placed in, which the
 * parity entry points (fmd_compute_forces, fmd_build_neighbors) use; equal to
 * fmd_get_box except under NPT between steps.  No reference counterpart. */
int32_t fmd_get_state_box(fmd_engine* e, int32_t sys, double* L);
/* Radius of gyration, averaged over chains (runs of equal molecule id in
 * global-id order), from unwrapped positions; replaces compute_rg
 * (src/analysis/rg.cuh:28-31).  per_chain (if non-NULL) receives up to cap
 * chain values; nchains (if non-NULL) the chain count. */
int32_t fmd_get_rg(fmd_engine* e, int32_t sys, double* rg, double* per_chain, int32_t cap,
                   int32_t* nchains);
/* fmd_get_rg with no device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The per-chain values and
 * their mean (summed on the device in chain order: the same bits as
 * fmd_get_rg) pass through one file-scope device scratch array per GPU of at
 * most FMD_RG_NOALLOC_MAX_CHAINS chains (512 KiB, held until the process
 * ends); a system with more chains gets FMD_ERR_ARG.  Calls that overlap on
 * one GPU share that array and race. */
#define FMD_RG_NOALLOC_MAX_CHAINS 65536
int32_t fmd_get_rg_noalloc(fmd_engine* e, int32_t sys, double* rg, double* per_chain, int32_t cap,
                           int32_t* nchains);
/* Exact checkpoint of a system's whole device state: positions, velocities
 * and images in the current sorted order, the Verlet list and its reference
 * positions, thermostat / barostat state, step and rebuild counters, the
 * stress correlator and unread thermo records.  A resumed system continues bit
 * for bit as if it had not stopped, in this or another engine of the same
 * library version.  The reference checkpoints through LAMMPS data restart
 * files only (src/main.cu:736-752), which keep positions, velocities and
 * images but not the thermostat or the list.  Call between steps. */
int32_t fmd_snapshot_size(fmd_engine* e, int32_t sys, uint64_t* bytes);
int32_t fmd_snapshot(fmd_engine* e, int32_t sys, void* buf, uint64_t cap);
int32_t fmd_system_resume(fmd_engine* e, const void* buf, uint64_t size, int32_t* sys);
/* The system's stress autocorrelator (a system created with stress_freq > 0).
 * fmd_corr_push adds one sample (xx, xy, xz, yy, yz, zz) as
 * MultipleTauCorrelator::push_sample (src/analysis/correlator.cu:94-99);
 * fmd_corr_reset clears it (correlator.cu:24-32); fmd_get_corr returns the
 * unnormalised sums corr[level][lag][6] and counts[level][lag], with
 * FMD_CORR_LEVELS x FMD_CORR_POINTS entries (get_results, correlator.cu:101-110).
 * Steps push the pressure tensor of every stress_freq-th step by themselves. */
int32_t fmd_corr_push(fmd_engine* e, int32_t sys, const double* sample6);
int32_t fmd_corr_reset(fmd_engine* e, int32_t sys);
int32_t fmd_get_corr(fmd_engine* e, int32_t sys, double* corr, int32_t* counts, int32_t* levels);
/* Latest reductions (thermo of the last completed step). */
int32_t fmd_get_thermo(fmd_engine* e, int32_t sys, fmd_thermo* out);
/* Drains the thermo records written every thermo_freq steps since the last
 * drain (oldest first); *n = number written; *lost = records overwritten. */
int32_t fmd_drain_thermo(fmd_engine* e, int32_t sys, fmd_thermo* out, int32_t cap, int32_t* n,
                         int64_t* lost);
/* Extended-system state of the Nose-Hoover chain and barostat (FP64), as
 * NoseHooverState holds it on the host (src/integrate/nose_hoover.cuh:14-45):
 * chain masses Q1 = Nf kT Tdamp^2 and Qr = kT Tdamp^2, barostat mass
 * W = (Nf + 3) kT Pdamp^2, V0, eps (V = V0 exp(3 eps)), v_eps, xi[k], v_xi[k]
 * (src/integrate/nose_hoover.cu:309-346).  Read between runs it is the state
 * after the last step's post-force chain / barostat update, which with the
 * velocities and energies makes up the conserved quantity of
 * tests/test_nose_hoover.cu:19-40.  ke_carry is the kinetic energy the next
 * pre-force chain application will use (chain_KE_carry). */
#define FMD_MAX_CHAIN 10
typedef struct fmd_thermostat {
    int32_t ensemble, M;
    int64_t step;
    double Q1, Qr, W, V0;
    double eps, veps;
    double xi[FMD_MAX_CHAIN], vxi[FMD_MAX_CHAIN];
    double ke_carry;
} fmd_thermostat;
int32_t fmd_get_thermostat(fmd_engine* e, int32_t sys, fmd_thermostat* out);

/* The device chain on a stream of kinetic energies: replaces
 * nh_propagate_chain_kernel (src/integrate/nose_hoover.cu:86-178) as its
 * GPUChainMatchesHost test drives it (tests/test_nose_hoover.cu:343-439).
 * Runs the step's FP64 chain code on the device, one draw after another with
 * the state carried: carry = 0 applies the chain once per draw with KE ke[t]
 * (scales[t]); carry = 1 applies it twice per draw, with ke[t] (the post-force
 * call, scales[2t]) and then with the KE that call leaves, ke[t] s^2 (the next
 * pre-force call on chain_KE_carry, scales[2t+1]).  xi and vxi (FMD_MAX_CHAIN
 * each) are the initial state in and the final state out. */
typedef struct fmd_chain_run {
    int32_t M, natoms, carry, n;
    double Q1, Qr, kT, half_dt;
    const double* ke;
} fmd_chain_run;
int32_t fmd_nh_chain_run(fmd_engine* e, const fmd_chain_run* c, double* xi, double* vxi, double* scales);

/* Neighbor-list rebuilds performed so far. */
int64_t fmd_system_rebuilds(fmd_engine* e, int32_t sys);

/* Times `reps` launches of the force kernel (terms as in fmd_compute_forces,
 * no kick, forces only: the work of a step that records no thermo) over the
 * listed systems with HIP events on the engine's stream; returns the average
 * launch time and the kernel's algorithmic HBM bytes per launch: 36 B per atom
 * (position, count, force) + 4 B per listed neighbor for LJ (whether the rows
 * are stored in 16 or 32 bits), + 4 B per atom header + 4 B per bond entry +
 * 8 B per angle entry for the bonded terms.  *pairs = listed neighbors.
 * Measurement only; no counterpart in the reference. */
int32_t fmd_time_force(fmd_engine* e, const int32_t* sys, int32_t nsys, uint32_t terms, int32_t reps,
                       double* avg_ms, double* alg_bytes, int64_t* pairs);

/* ----- autotuner -------------------------------------------------------- */
/* Picks, per replica size class, the force pass's workgroup shape (64, 128 or
 * 256 threads) and neighbor layout (the list build writes rows as offsets into
 * each group's neighbor set staged in LDS, or as slots gathered from HBM),
 * timing every candidate on copies of the given systems (force pass plus a
 * tenth of the list build in that layout).  The systems themselves are not
 * touched; every candidate computes the same bits.  The step then launches the
 * force pass per class range with that class's shape; a new layout reaches a
 * system at its next rebuild.  No reference counterpart (the reference
 * hard-codes 256-thread blocks, src/force/lj.cu:116-118). */
int32_t fmd_autotune(fmd_engine* e, const int32_t* sys, int32_t nsys, int32_t reps);
/* Chosen value for (kernel, size class): kernel 0 = force block shape,
 * 1 = list build block shape (always 256: one group of four tiles),
 * 2 = neighbor layout (1 = staged set, 0 = direct slots). */
int32_t fmd_tuned_block(const fmd_engine* e, int32_t kernel, int32_t size_class);
/* Sets one tuned value (as fmd_tuned_block names them) by hand: the same
 * bits whatever the value; a layout reaches a system at its next rebuild. */
int32_t fmd_engine_set_tuned(fmd_engine* e, int32_t kernel, int32_t size_class, int32_t value);
/* The last autotune's timings of a size class: up to cap rows of (layout,
 * block, average force-pass ms); returns the row count. */
int32_t fmd_autotune_times(const fmd_engine* e, int32_t size_class, int32_t* layout, int32_t* block, double* ms,
                           int32_t cap);
/* Size class of a system with natoms beads (0: < 8192, 1: < 32768, 2: rest). */
int32_t fmd_size_class(int32_t natoms);

/* ----- serving: continuous batching under an HBM budget ----------------- */
typedef struct fmd_server_opts {
    uint64_t hbm_budget;       /* bytes the server's resident jobs may hold */
    int32_t max_jobs;          /* resident job cap (0 = 256)                 */
    int32_t graph_steps;       /* steps per batched graph launch (0 = 20)    */
} fmd_server_opts;
typedef struct fmd_job_result {
    int32_t status;            /* FMD_OK or an error code                    */
    int32_t natoms;
    int64_t steps;
    double t_submit, t_admit, t_done;  /* seconds on the server clock        */
    fmd_thermo last;
} fmd_job_result;
enum { FMD_JOB_QUEUED = 0, FMD_JOB_RUNNING = 1, FMD_JOB_DONE = 2, FMD_JOB_FAILED = 3 };

int32_t fmd_server_create(int32_t device, const fmd_server_opts* o, fmd_server** out);
/* Queues a job (copies the descriptor); it runs nsteps steps. */
int32_t fmd_server_submit(fmd_server* s, const fmd_system_desc* d, int64_t nsteps, int64_t* job);
int32_t fmd_server_poll(fmd_server* s, int64_t job, int32_t* state);
/* Result and, if non-NULL, final state by global id (pos4, vel3). */
int32_t fmd_server_result(fmd_server* s, int64_t job, fmd_job_result* r, float* pos4,
                          float* vel3);
/* Blocks until every submitted job is done or failed. */
int32_t fmd_server_drain(fmd_server* s);
/* Server clock (seconds since create), bead-steps done, steps of the batch. */
double fmd_server_now(const fmd_server* s);
int32_t fmd_server_stats(const fmd_server* s, double* bead_steps, int64_t* graph_launches,
                         int32_t* resident, int32_t* queued);
void fmd_server_destroy(fmd_server* s);
/* Tuned value the server's engine uses for (kernel 0 = force block shape,
 * 1 = list build block shape, 2 = neighbor layout; size class), e.g. copied
 * from fmd_tuned_block of an autotuned engine; takes effect at the next
 * batched launch.  No reference counterpart (the reference
 * hard-codes 256-thread blocks, src/force/lj.cu:116-118). */
int32_t fmd_server_set_block(fmd_server* s, int32_t kernel, int32_t size_class, int32_t block);

/* Pure admission policy (no GPU): replays arrivals (t, bytes, duration) under a
 * byte budget and a job cap with FIFO admission when memory frees, as the
 * server does; writes each job's admit and finish times. */
int32_t fmd_admission_replay(int32_t njobs, const double* t_arrive, const uint64_t* bytes,
                             const double* duration, uint64_t budget, int32_t max_jobs,
                             double* t_admit, double* t_finish);

#ifdef __cplusplus
}
#endif
#endif /* FMD_FMD_H */
