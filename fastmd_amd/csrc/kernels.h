// kernels.h -- host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fmd_device.h"

namespace fmd {

// Launch geometry shared by the launchers.
struct Launch {
    const RepDesc* descs;  // device table
    const Tile* tiles;     // device tile list
    int ntl;               // tiles in the list
    const int* reps;       // device list of replica indices
    int nr;                // replicas in the list
    const int* rep_t0;     // first tile of each listed replica
    int* rlist;            // compact list of rebuilding tiles (capacity ntl)
    int* nrt;              // its length
    const Tile* chunks;    // (replica, chunk) list for the first reduction level
    int nch;
    hipStream_t stream;
};

enum BoundaryMode { BND_PRIME = 0, BND_STEP = 1, BND_START = 2, BND_REDUCE = 3 };
enum ControlOp { CTL_RUN = 0, CTL_FORCE_REBUILD = 1 };

void launch_pre(const Launch& l);
void launch_decide(const Launch& l);
// The rebuild phases (Morton cell sort and, if `build`, the Verlet lists) of
// the tiles in the rebuild tile list: every tile if `always` (listed here),
// else those k_decide listed this step (the replicas it found rebuilding).
void launch_rebuild(const Launch& l, int always, int build);
// The Verlet lists of the tiles the last rebuild phases listed, by at most
// max_blocks workgroups (0: the default cap), each taking listed groups in turn.
void launch_build(const Launch& l, int max_blocks = 0);
// Force kernel over tiles [t0, t0+nt) of the list with the given block size.
// Force kernel modes: FM_STEP skips inactive replicas and records PE and
// virial on the steps that want them; FM_OBS always runs and records them;
// FM_FAST always runs and records forces only (a plain step's work).
enum ForceMode { FM_STEP = 0, FM_OBS = 1, FM_FAST = 2 };
// Tile selection: every tile of the range, or only those of replicas that
// keep their lists this step (FS_STAYING).
enum ForceSel { FS_ALL = 0, FS_STAYING = 1 };
void launch_force(const Launch& l, int t0, int nt, int block, int terms, int kick, int mode, int sel = FS_ALL);
// All terms over the tiles of the replicas that rebuilt their lists this step
// (the rebuild phases' tile list), a workgroup per group of tiles.
void launch_force_listed(const Launch& l, int kick, int mode);
void launch_boundary(const Launch& l, int mode);
void launch_control(const Launch& l, int op, int64_t arg);
// Radius of gyration of each of a replica's chains.
void launch_rg(const RepDesc* d, int nchains, double* out, hipStream_t s);
// Their mean (the replica's chain count of values in vals) into out[0].
void launch_rg_mean(const RepDesc* d, const double* vals, double* out, hipStream_t s);
// One stress sample (xx, xy, xz, yy, yz, zz), passed by value.
struct Sample6 { double v[6]; };
void launch_corr_push(const RepDesc* d, const Sample6& x, hipStream_t s);
// Chain applications of nh_chain on a stream of kinetic energies (one lane).
void launch_nh_chain_run(RepState* st, const double* ke, int n, int carry, double kT, int natoms,
                         double half_dt, double* scales, hipStream_t s);
// Per replica of the list: (err, active, step lo, step hi).
void launch_status(const Launch& l, int4* out);
// Per replica of the list: steps since its last list build (-1: inactive).
void launch_phase(const Launch& l, int* out);
// Reference cell ids / Morton keys by global id for one replica.
void launch_ref_cells(const RepDesc* d, int natoms, int* out, hipStream_t s);
void launch_morton_keys(const RepDesc* d, int natoms, uint32_t* out, hipStream_t s);

}  // namespace fmd
