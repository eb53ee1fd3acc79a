// engine.h -- the device-resident multi-replica MD engine (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "fmd/fmd.h"
#include "fmd_device.h"
#include "kernels.h"

namespace fmd {

constexpr int kNumClasses = 3;
constexpr int kRgScratchChains = 1 << 16;  // chains the allocation-free Rg read holds
constexpr int kGraphSteps = 20;
// batches of at least this many replicas, with their list builds spread over
// at least kForkPhases of the kRebuildEvery steps, build lists beside the
// force pass
constexpr int kForkReplicas = 8;
constexpr int kForkPhases = kRebuildEvery / 2;
constexpr int64_t kForkRecheck = 50 * kRebuildEvery;  // steps between two choices by a waiting caller
// share of the staying replicas' force pass launched beside the cell sort,
// before the list build is queued (measured on 255 staggered replicas: 0.15,
// 0.3, 0.5, 0.6, 0.7 and 0.85 gave 0.737, 0.730, 0.716, 0.723, 0.738 and
// 0.774 ms per step, 0.740 without this part)
constexpr double kEarlyForce = 0.5;
// workgroups of the list build in that step, the GPU's slots for it (256 CUs x
// 4): the rebuilding replicas hold about a twentieth of the groups, and
// workgroups past the list only take slots beside the force pass (measured on
// 255 staggered replicas against 8,192: 4,096 -3.1 us, 2,048 -4.7 to -6.5 us,
// 1,024 -5.8 us per step)
constexpr int kForkBuildBlocks = 1024;

int size_class(int natoms);

struct Replica {
    int id = -1;
    int natoms = 0, nslots = 0, ntiles = 0, cls = 0;
    int ensemble = 0;
    RepDesc desc{};         // host copy of the device descriptor
    RepState* st = nullptr; // device
    void* block = nullptr;  // the replica's device arrays, carved from one block
    size_t block_bytes = 0;
    size_t block_used = 0;  // bytes of the block the arrays occupy
    uint64_t bytes = 0;
    int64_t steps_done = 0; // host mirror of st->step
    int64_t thermo_read = 0;
    int thermo_cap = 0;
};

struct Batch {
    std::vector<int> ids;   // replica ids, ordered by size class
    Tile* d_tiles = nullptr;
    int* d_reps = nullptr;
    int* d_rep_t0 = nullptr;  // first tile of each replica in d_reps order
    int* d_rlist = nullptr;   // compact rebuild tile list (ntl entries)
    int* d_nrt = nullptr;     // its length
    Tile* d_chunks = nullptr; // (replica, chunk of kChunkTiles tiles) list
    int nch = 0;
    int4* d_status = nullptr; // per replica (err, active, step lo, step hi), d_reps order
    int* d_phase = nullptr;   // per replica, steps since its last list build (d_reps order)
    bool fork = false;        // the step builds lists beside the force pass (launch_step_seq)
    bool fork_known = false;
    int64_t steps_since_fork = 0;  // steps launched since that choice
    std::vector<int> ord;     // host copy of d_reps
    void* block = nullptr;    // the arrays above, carved from one device block
    size_t block_bytes = 0;
    int ntl = 0, nr = 0;
    int cls_t0[kNumClasses] = {0, 0, 0}, cls_nt[kNumClasses] = {0, 0, 0};
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    int tune_version = -1;
};

// Options of Engine::create beyond the descriptor.
struct CreateOpts {
    int kmult0 = 1;          // first neighbor-capacity multiplier to try (1, 2 or 4)
    bool side = false;       // uploads and the initial build on the side stream
    // out
    uint64_t charged = 0;    // bytes charged against the budget (success)
    uint64_t need = 0;       // bytes the last attempt needed (FMD_ERR_BUDGET)
    int kmult = 1;           // multiplier of the last attempt
};

// Device bytes a system is charged: the block its arrays are carved from,
// in whole 2 MiB allocation granules.
uint64_t charge_bytes(uint64_t block_bytes);

class Engine {
  public:
    Engine(int device, uint64_t budget);
    ~Engine();
    bool ok() const { return init_ok_; }
    const char* last_error() const { return err_.c_str(); }
    int32_t fail(int32_t code, const std::string& msg);

    int32_t create(const fmd_system_desc* d, int32_t* sys, CreateOpts* o = nullptr);
    // wait = false: the caller guarantees no launch in flight uses the system
    // (it is in no batch of work still queued on the main stream).
    int32_t destroy(int32_t sys, bool wait = true);
    // While deferred, blocks released to the cache are not returned to HIP
    // (hipFree waits for the whole device); flush_deferred() trims the cache.
    void set_defer_free(bool on) { defer_free_ = on; }
    void flush_deferred();
    uint64_t bytes_in_use() const { return in_use_; }
    uint64_t budget() const { return budget_; }

    int32_t set_force_cap(int32_t sys, float cap);
    int32_t set_langevin(int32_t sys, float T0, float T1, float Tdamp);
    int32_t build_neighbors(int32_t sys);
    int32_t get_pairs(int32_t sys, int64_t* n, int32_t* ij, int64_t cap);
    int32_t get_cell_ids(int32_t sys, int32_t* out, int32_t* nx);
    int32_t get_morton(int32_t sys, uint32_t* out);
    int32_t compute_forces(int32_t sys, uint32_t terms, float* f4, double* vir, double* pe);

    int32_t step(const int32_t* ids, int32_t n, int64_t nsteps, bool use_graph);
    // Non-blocking pieces used by the server: set run lengths, launch k steps.
    int32_t begin_run(const std::vector<int>& ids, int64_t nsteps);
    // The same on the side stream, completed before it returns (for systems
    // created there while the main stream is busy).
    int32_t begin_run_side(const std::vector<int>& ids, int64_t nsteps);
    int32_t launch_steps(const std::vector<int>& ids, int64_t k, bool use_graph, bool may_sync = false);
    int32_t sync();
    int32_t check_errors(const std::vector<int>& ids, std::vector<int>* failed);
    int32_t replica_error(int32_t sys);
    // Frees every cached batch except the one for `keep` (sorted or not).
    void trim_batches(const std::vector<int>& keep);

    int64_t step_count(int32_t sys);
    int64_t rebuilds(int32_t sys);
    // wait = false: read without waiting for the main stream (the system is
    // not in any launch still in flight)
    int32_t get_state(int32_t sys, float* pos4, float* vel3, int32_t* img3, bool wait = true);
    int32_t get_box(int32_t sys, double* L, bool of_positions = false);
    int32_t get_thermo(int32_t sys, fmd_thermo* out);
    int32_t get_thermostat(int32_t sys, fmd_thermostat* out);
    int32_t nh_chain_run(const fmd_chain_run* c, double* xi, double* vxi, double* scales);
    int32_t get_rg(int32_t sys, double* rg, double* per_chain, int32_t cap, int32_t* nchains);
    // The same with no device allocation: the per-chain values and their mean
    // go through the device's file-scope scratch array (at most kRgScratchChains).
    int32_t get_rg_noalloc(int32_t sys, double* rg, double* per_chain, int32_t cap, int32_t* nchains);
    int32_t snapshot_size(int32_t sys, uint64_t* bytes);
    int32_t snapshot(int32_t sys, void* buf, uint64_t cap);
    int32_t resume(const void* buf, uint64_t size, int32_t* sys);
    int32_t corr_push(int32_t sys, const double* x);
    int32_t corr_reset(int32_t sys);
    int32_t get_corr(int32_t sys, double* corr, int32_t* counts, int32_t* levels);
    int32_t drain_thermo(int32_t sys, fmd_thermo* out, int32_t cap, int32_t* n, int64_t* lost);

    int32_t autotune(const int32_t* ids, int32_t n, int reps);
    int32_t time_force(const int32_t* ids, int32_t n, uint32_t terms, int reps, double* ms,
                       double* bytes, int64_t* pairs);
    int tuned(int kernel, int cls) const { return tuned_[kernel][cls]; }
    // Sets one tuned parameter of a size class: kernel 0 = force block shape
    // (64, 128 or 256), 1 = list build (one shape, kGroupSlots), 2 = the force
    // pass's neighbor layout (1 = staged neighbor set, 0 = direct gathers).
    int32_t set_tuned(int kernel, int cls, int value);
    // Timings of the last autotune: per size class, (layout, block, ms) rows.
    int32_t tune_times(int cls, int32_t* layout, int32_t* block, double* ms, int32_t cap) const;
    hipStream_t stream() const { return stream_; }
    Replica* rep(int32_t sys);

  private:
    int32_t create_impl(const fmd_system_desc* d, int32_t* sys, int kmult, bool side, uint64_t* charged,
                        uint64_t* need);
    int32_t hip_check(hipError_t e, const char* what);
    void* acquire_block(size_t bytes, size_t* got);
    void release_block(void* p, size_t bytes);
    void flush_cache(uint64_t keep);
    void free_batch(Batch* b);
    Batch* batch_for(const std::vector<int>& ids);
    void drop_batches_with(int id);
    void launch_step_seq(Batch& b);
    void decide_fork(Batch& b, bool may_sync);
    Launch launch_of(const Batch& b) const;
    Launch launch_of(const Batch& b, hipStream_t s) const;
    int32_t read_state(Replica& r, RepState& s);

    int device_ = 0;
    bool init_ok_ = false;
    std::string err_;
    hipStream_t stream_ = nullptr;
    hipStream_t side_ = nullptr;  // system creation beside the main stream's launches
    hipStream_t aux_ = nullptr;   // a step's list rebuilds beside its force pass
    hipEvent_t ev_fork_ = nullptr, ev_sorted_ = nullptr, ev_join_ = nullptr;
    double* scratch_ = nullptr;   // this device's scratch array (device address), or null
    bool defer_free_ = false;
    uint64_t budget_ = 0, in_use_ = 0;
    std::vector<Replica*> reps_;
    RepDesc* d_descs_ = nullptr;
    std::map<std::vector<int>, Batch*> batches_;
    int tuned_[3][kNumClasses];
    int tune_version_ = 0;
    struct TuneRow { int layout, block; double ms; };
    std::vector<TuneRow> tune_rows_[kNumClasses];
    void apply_layout(int cls);  // the class's neighbor layout into its systems' descriptors
    std::multimap<size_t, void*> cache_;  // released replica blocks by size
    uint64_t cached_bytes_ = 0;
};

}  // namespace fmd
