// engine.hip -- host side of the device-resident multi-replica MD engine.
//
// Each replica owns its device buffers (sized at creation, never resized) and
// one RepDesc entry in a device table of kMaxReplicas.  A Batch is a set of
// replicas with its device tile list; its K-step launch sequence is captured
// once into a hipGraph and replayed.  Nothing in the step loop reads back to
// the host: rebuild decisions, thermostat/barostat scalars and thermo records
// all stay on the device.
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "kernels.h"

namespace fmd {

int size_class(int natoms) { return natoms < 8192 ? 0 : (natoms < 32768 ? 1 : 2); }

// slot count: whole groups of kGroupTiles tiles
static inline int ceil_slots(int n) { return (n + kGroupSlots - 1) / kGroupSlots * kGroupSlots; }


// Device scratch of the allocation-free Rg read: the per-chain values, then
// their mean.  One copy per GPU, at a device address found once per device
// (hipGetSymbolAddress with that device current); host code uses that address,
// never the symbol.  Calls that overlap on one GPU share it and race.
static_assert(kRgScratchChains == FMD_RG_NOALLOC_MAX_CHAINS, "the header states the scratch limit");
static __device__ __align__(256) double scratch[kRgScratchChains + 32];
static constexpr int kMaxScratchDevices = 64;
static double* g_scratch_ptr[kMaxScratchDevices];
static std::mutex g_scratch_mu;

Engine::Engine(int device, uint64_t budget) : device_(device) {
    for (int c = 0; c < kNumClasses; c++) {
        tuned_[0][c] = 256;
        tuned_[1][c] = kGroupSlots;  // the list build runs one workgroup per group of tiles
        tuned_[2][c] = 1;            // staged neighbor sets
    }
    if (hip_check(hipSetDevice(device), "hipSetDevice")) return;
    if (hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate")) return;
    if (hip_check(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "hipStreamCreate")) return;
    if (hip_check(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking), "hipStreamCreate")) return;
    if (hip_check(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming), "hipEventCreate")) return;
    if (hip_check(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming), "hipEventCreate")) return;
    if (hip_check(hipEventCreateWithFlags(&ev_sorted_, hipEventDisableTiming), "hipEventCreate")) return;
    size_t fr = 0, tot = 0;
    if (hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo")) return;
    budget_ = budget ? budget : (uint64_t)(0.9 * (double)fr);
    if (hip_check(hipMalloc(&d_descs_, sizeof(RepDesc) * kMaxReplicas), "hipMalloc(descs)")) return;
    reps_.assign(kMaxReplicas, nullptr);
    if (device >= 0 && device < kMaxScratchDevices) {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        if (!g_scratch_ptr[device]) {
            void* p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess && p)
                g_scratch_ptr[device] = (double*)p;
            else
                (void)hipGetLastError();  // the allocation-free read reports it; the rest works
        }
        scratch_ = g_scratch_ptr[device];
    }
    init_ok_ = true;
}

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (side_) (void)hipStreamSynchronize(side_);
    if (aux_) (void)hipStreamSynchronize(aux_);
    for (auto& kv : batches_) {
        free_batch(kv.second);
    }
    for (Replica* r : reps_)
        if (r) {
            (void)hipFree(r->block);
            delete r;
        }
    flush_cache(0);
    if (d_descs_) (void)hipFree(d_descs_);
    if (stream_) (void)hipStreamDestroy(stream_);
    if (side_) (void)hipStreamDestroy(side_);
    if (aux_) (void)hipStreamDestroy(aux_);
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (ev_sorted_) (void)hipEventDestroy(ev_sorted_);
}

int32_t Engine::fail(int32_t code, const std::string& msg) {
    err_ = msg;
    return code;
}

int32_t Engine::hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return FMD_OK;
    err_ = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? FMD_ERR_NOMEM : FMD_ERR_HIP;
}

Replica* Engine::rep(int32_t sys) {
    if (sys < 0 || sys >= kMaxReplicas) return nullptr;
    return reps_[sys];
}

// Replica blocks are cached when a replica is destroyed and reused for a new
// one of similar size, so serving does not call hipMalloc / hipFree per job.
void* Engine::acquire_block(size_t bytes, size_t* got) {
    auto it = cache_.lower_bound(bytes);
    if (it != cache_.end() && it->first <= bytes + bytes / 4) {
        void* p = it->second;
        *got = it->first;
        cached_bytes_ -= it->first;
        cache_.erase(it);
        return p;
    }
    const size_t want = (bytes + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);  // 2 MiB granules
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        flush_cache(0);
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    *got = want;
    return p;
}

void Engine::release_block(void* p, size_t bytes) {
    if (!p) return;
    cache_.emplace(bytes, p);
    cached_bytes_ += bytes;
    flush_cache(std::min<uint64_t>(budget_ / 8, 8ull << 30));
}

void Engine::flush_deferred() {
    defer_free_ = false;
    flush_cache(std::min<uint64_t>(budget_ / 8, 8ull << 30));
}

void Engine::flush_cache(uint64_t keep) {
    if (defer_free_ && keep > 0) return;
    while (cached_bytes_ > keep && !cache_.empty()) {
        auto it = std::prev(cache_.end());  // largest first
        (void)hipFree(it->second);
        cached_bytes_ -= it->first;
        cache_.erase(it);
    }
}

// ---------------------------------------------------------------------------
// sizing
// ---------------------------------------------------------------------------
struct Sizing {
    int natoms, nslots, ntiles, kcap, bcap, acap, ncell_cap, thermo_cap;
    uint64_t bytes;
};

static int cbits_for(int nx) {
    int cb = 0;
    while ((1 << cb) < nx) cb++;
    return cb;
}

static Sizing sizing(const fmd_system_desc* d, int bcap, int acap, int kmult = 1) {
    Sizing s;
    s.natoms = d->natoms;
    s.nslots = ceil_slots(d->natoms);
    s.ntiles = s.nslots / 64;
    const double rcs = (double)d->rc + (double)d->skin;
    const double rho = (double)d->natoms / ((double)d->L * d->L * d->L);
    // capacity rule of src/neighbor/verlet_list.cu:4-7: max(256, 3 rho 4/3 pi rcs^3);
    // only the filled rows are streamed, so the reserve costs memory, not bandwidth
    const double expect = rho * 4.0 / 3.0 * M_PI * rcs * rcs * rcs;
    int kc = std::max(256, (int)(3.0 * expect)) * kmult;
    kc = (kc + 7) / 8 * 8;
    if (kc > s.nslots) kc = s.nslots;  // cannot have more than N-1 neighbors
    if (kc < 8) kc = 8;
    s.kcap = kc;
    s.bcap = std::max(1, bcap);
    s.acap = std::max(1, acap);
    float rcsf = d->rc + d->skin;
    double growth = (d->ensemble == FMD_NPT_NH) ? 1.5 : 1.0;
    int nxmax = std::max(1, (int)(growth * (double)d->L / (double)rcsf) + 1);
    int cb = cbits_for(nxmax);
    s.ncell_cap = 1 << (3 * cb);
    s.thermo_cap = d->thermo_freq > 0 ? 4096 : 0;
    uint64_t b = 0;
    b += (uint64_t)s.nslots * (9 * 16 + 8 * 4);
    b += (uint64_t)s.natoms * 4;
    b += (uint64_t)s.ncell_cap * 8 + 4;
    b += (uint64_t)s.ntiles * 64 * ((uint64_t)s.kcap * 4 + (uint64_t)s.bcap * 4 + (uint64_t)s.acap * 8);
    b += (uint64_t)(s.ntiles / kGroupTiles) * (4 + (uint64_t)kGroupStage * 4) + (uint64_t)s.nslots + 768;
    b += (uint64_t)s.ntiles * (kNumPart * 8 + 4 + 4) + (uint64_t)((s.ntiles + kChunkTiles - 1) / kChunkTiles) * kNumPart * 8;
    b += (uint64_t)(d->natoms + 1) * 8 + (uint64_t)d->nbonds * 8 + (uint64_t)d->nangles * 48;
    b += sizeof(RepState) + (uint64_t)s.thermo_cap * sizeof(ThermoRec) + 4096;
    for (int k = 0; k < d->ntables; k++) b += 24 + (uint64_t)d->table_n[k] * 8 + 1024;
    if (d->stress_freq > 0) b += sizeof(CorrState) + 256;
    s.bytes = b;
    return s;
}

// Per-atom bonded incidence counts (max over atoms), needed for the table caps.
static void incidence_caps(const fmd_system_desc* d, int* bcap, int* acap) {
    std::vector<int> nb(d->natoms, 0), na(d->natoms, 0);
    for (int b = 0; b < d->nbonds; b++) { nb[d->bonds[2 * b]]++; nb[d->bonds[2 * b + 1]]++; }
    for (int a = 0; a < d->nangles; a++)
        for (int q = 0; q < 3; q++) na[d->angles[4 * a + q]]++;
    *bcap = 0; *acap = 0;
    for (int i = 0; i < d->natoms; i++) { *bcap = std::max(*bcap, nb[i]); *acap = std::max(*acap, na[i]); }
}

static int32_t validate(const fmd_system_desc* d, std::string& msg) {
    if (!d) { msg = "null descriptor"; return FMD_ERR_ARG; }
    if (d->natoms <= 0 || d->natoms > kSlotMask) { msg = "natoms out of range"; return FMD_ERR_ARG; }
    if (d->ntypes < 1 || d->ntypes > kMaxTypes) { msg = "ntypes must be in [1,16]"; return FMD_ERR_ARG; }
    if (!d->pos || !d->lj) { msg = "pos and lj are required"; return FMD_ERR_ARG; }
    if (!(d->L > 0.0f) || !(d->rc > 0.0f) || d->skin < 0.0f || !(d->dt > 0.0f)) {
        msg = "need L > 0, rc > 0, skin >= 0, dt > 0";
        return FMD_ERR_ARG;
    }
    if (d->ensemble < 0 || d->ensemble > 2) { msg = "bad ensemble"; return FMD_ERR_ARG; }
    if (!(d->Tdamp > 0.0f)) { msg = "Tdamp must be > 0"; return FMD_ERR_ARG; }
    if (d->ensemble == FMD_NPT_NH && !(d->Pdamp > 0.0f)) { msg = "Pdamp must be > 0"; return FMD_ERR_ARG; }
    if (d->ensemble != FMD_LANGEVIN && (d->chain_length < 1)) { msg = "chain_length must be >= 1"; return FMD_ERR_ARG; }
    for (int i = 0; i < d->natoms; i++) {
        int t = type_of(d->pos[4 * i + 3]);
        if (t < 0 || t >= d->ntypes) { msg = "atom type out of range"; return FMD_ERR_ARG; }
    }
    if (d->nbonds < 0 || d->nangles < 0) { msg = "negative counts"; return FMD_ERR_ARG; }
    if (d->stress_freq < 0) { msg = "stress_freq must be >= 0"; return FMD_ERR_ARG; }
    if (d->ntables < 0 || d->ntables > 4096) { msg = "ntables out of range"; return FMD_ERR_ARG; }
    if (d->ntables > 0) {
        if (!d->table_idx || !d->table_n || !d->table_range || !d->table_data) {
            msg = "tables need table_idx, table_n, table_range, table_data";
            return FMD_ERR_ARG;
        }
        for (int k = 0; k < d->ntables; k++) {
            if (d->table_n[k] < 2 || d->table_n[k] > (1 << 24)) { msg = "a table needs 2 to 2^24 points"; return FMD_ERR_ARG; }
            const float lo = d->table_range[2 * k], hi = d->table_range[2 * k + 1];
            if (!(hi > lo) || !std::isfinite(lo) || !std::isfinite(hi)) { msg = "table range must have rmax > rmin"; return FMD_ERR_ARG; }
        }
        for (int p = 0; p < d->ntypes * d->ntypes; p++) {
            const int ti = p / d->ntypes, tj = p % d->ntypes, k = d->table_idx[p];
            if (k < -1 || k >= d->ntables) { msg = "table index out of range"; return FMD_ERR_ARG; }
            if (k != d->table_idx[tj * d->ntypes + ti]) { msg = "table_idx must be symmetric"; return FMD_ERR_ARG; }
            if (k >= 0 && (d->lj[2 * p] != 0.0f || d->lj[2 * p + 1] != 0.0f)) {
                msg = "both lj and table defined for a type pair";
                return FMD_ERR_ARG;
            }
        }
    }
    if (d->nbonds > 0 && (!d->bonds || !d->bond_types || !d->bond_params)) { msg = "bonds need bonds, bond_types, bond_params"; return FMD_ERR_ARG; }
    if (d->nangles > 0 && (!d->angles || !d->angle_params)) { msg = "angles need angles and angle_params"; return FMD_ERR_ARG; }
    for (int b = 0; b < d->nbonds; b++) {
        int i = d->bonds[2 * b], j = d->bonds[2 * b + 1], t = d->bond_types[b];
        if (i < 0 || j < 0 || i >= d->natoms || j >= d->natoms || i == j) { msg = "bond atom index out of range"; return FMD_ERR_ARG; }
        if (t < 0 || t >= d->nbond_types || t > 63) { msg = "bond type out of range"; return FMD_ERR_ARG; }
    }
    for (int a = 0; a < d->nangles; a++) {
        for (int q = 0; q < 3; q++) {
            int i = d->angles[4 * a + q];
            if (i < 0 || i >= d->natoms) { msg = "angle atom index out of range"; return FMD_ERR_ARG; }
        }
        int t = d->angles[4 * a + 3];
        if (t < 0 || t >= d->nangle_types || t > 63) { msg = "angle type out of range"; return FMD_ERR_ARG; }
    }
    return FMD_OK;
}

uint64_t charge_bytes(uint64_t block_bytes) {
    const uint64_t g = 2ull << 20;
    return (block_bytes + g - 1) / g * g;
}

// The admission unit: what create charges at the first neighbor capacity.
uint64_t system_bytes(const fmd_system_desc* d) {
    std::string m;
    if (validate(d, m)) return 0;
    int bc, ac;
    incidence_caps(d, &bc, &ac);
    return charge_bytes(sizing(d, bc, ac).bytes);
}

// ---------------------------------------------------------------------------
// create / destroy
// ---------------------------------------------------------------------------
int32_t Engine::create(const fmd_system_desc* d, int32_t* sys, CreateOpts* o) {
    // an initial configuration denser than the list capacity rule allows is
    // created again with a larger capacity instead of being truncated
    CreateOpts dflt;
    if (!o) o = &dflt;
    int32_t rc = FMD_ERR_NBR_OVERFLOW;
    int kmult = o->kmult0 == 2 || o->kmult0 == 4 ? o->kmult0 : 1;
    for (; kmult <= 4 && rc == FMD_ERR_NBR_OVERFLOW; kmult *= 2) {
        o->kmult = kmult;
        rc = create_impl(d, sys, kmult, o->side, &o->charged, &o->need);
    }
    return rc;
}

int32_t Engine::create_impl(const fmd_system_desc* d, int32_t* sys, int kmult, bool side, uint64_t* charged,
                            uint64_t* need) {
    std::string msg;
    if (int32_t rc = validate(d, msg)) return fail(rc, msg);
    int bcap, acap;
    incidence_caps(d, &bcap, &acap);
    if (bcap > 255 || acap > 255) return fail(FMD_ERR_ARG, "more than 255 bonds or angles on one atom");
    Sizing S = sizing(d, bcap, acap, kmult);
    uint64_t charge = charge_bytes(S.bytes);
    *need = charge;
    if (in_use_ + charge > budget_)
        return fail(FMD_ERR_BUDGET, "system needs " + std::to_string(charge) +
                                        " bytes; budget left " + std::to_string(budget_ - in_use_));
    int id = -1;
    for (int k = 0; k < kMaxReplicas; k++)
        if (!reps_[k]) { id = k; break; }
    if (id < 0) return fail(FMD_ERR_BUDGET, "replica table full");
    Replica* R = new Replica();
    Replica& r = *R;
    r.id = id;
    r.natoms = S.natoms; r.nslots = S.nslots; r.ntiles = S.ntiles;
    r.cls = size_class(S.natoms);
    r.ensemble = d->ensemble;
    r.thermo_cap = S.thermo_cap;
    RepDesc& D = r.desc;
    std::memset(&D, 0, sizeof(D));
    D.natoms = S.natoms; D.nslots = S.nslots; D.ntiles = S.ntiles; D.ntypes = d->ntypes;
    D.kcap = S.kcap; D.bcap = S.bcap; D.acap = S.acap; D.ncell_cap = S.ncell_cap;
    D.ensemble = d->ensemble;
    D.rc2 = d->rc * d->rc;
    D.rcs = d->rc + d->skin;
    D.rcs2 = D.rcs * D.rcs;
    float hs = d->skin * 0.5f;
    D.half_skin2 = hs * hs;
    D.dt = d->dt;
    D.half_dt = 0.5f * d->dt;
    D.c1 = std::exp(-(1.0f / d->Tdamp) * d->dt);  // src/integrate/langevin.cu:17-19
    D.c2 = std::sqrt(1.0f - D.c1 * D.c1);
    D.fcap = d->force_cap > 0.0f ? d->force_cap : 0.0f;
    D.vlim = D.fcap > 0.0f ? 0.02f / d->dt : 0.0f;
    D.clamp_lj = 1;
    if (d->ntypes == 1 && d->lj[0] != 0.0f && d->lj[1] != 0.0f) {  // the force pass's one-type form
        D.lj_kf = 24.0f * d->lj[0] / (d->lj[1] * d->lj[1]);
        D.lj_fclamp = 1000.0f / D.lj_kf;
    }
    D.idx16 = (S.nslots <= 65536) && std::getenv("FMD_IDX32") == nullptr ? 1 : 0;
    D.nostage = tuned_[2][size_class(S.natoms)] ? 0 : 1;
    D.seed_lo = (uint32_t)d->seed;
    D.seed_hi = (uint32_t)(d->seed >> 32);

    const size_t ns = S.nslots, nt = S.ntiles, na = S.natoms;
    // topology by global id
    std::vector<int> boff(na + 1, 0), aoff(na + 1, 0);
    for (int b = 0; b < d->nbonds; b++) { boff[d->bonds[2 * b] + 1]++; boff[d->bonds[2 * b + 1] + 1]++; }
    for (int a = 0; a < d->nangles; a++)
        for (int q = 0; q < 3; q++) aoff[d->angles[4 * a + q] + 1]++;
    for (size_t i = 0; i < na; i++) { boff[i + 1] += boff[i]; aoff[i + 1] += aoff[i]; }
    std::vector<int> blst(std::max(1, boff[na]));
    std::vector<int4> alst(std::max(1, aoff[na]));
    {
        std::vector<int> fb(boff.begin(), boff.end() - 1), fa(aoff.begin(), aoff.end() - 1);
        for (int b = 0; b < d->nbonds; b++) {
            int i = d->bonds[2 * b], j = d->bonds[2 * b + 1], t = d->bond_types[b];
            blst[fb[i]++] = j | (t << kSlotBits);
            blst[fb[j]++] = i | (t << kSlotBits);
        }
        for (int a = 0; a < d->nangles; a++) {
            int4 A = make_int4(d->angles[4 * a], d->angles[4 * a + 1], d->angles[4 * a + 2], d->angles[4 * a + 3]);
            alst[fa[A.x]++] = A;
            if (A.y != A.x) alst[fa[A.y]++] = A;
            if (A.z != A.x && A.z != A.y) alst[fa[A.z]++] = A;
        }
        // an atom listed twice in one angle gets one entry; fix the offsets
        std::vector<int> cnt(na, 0);
        for (int a = 0; a < d->nangles; a++) {
            int x = d->angles[4 * a], y = d->angles[4 * a + 1], z = d->angles[4 * a + 2];
            cnt[x]++;
            if (y != x) cnt[y]++;
            if (z != x && z != y) cnt[z]++;
        }
        std::vector<int> aoff2(na + 1, 0);
        std::vector<int4> alst2(std::max(1, aoff[na]));
        for (size_t i = 0; i < na; i++) {
            aoff2[i + 1] = aoff2[i] + cnt[i];
            for (int k = 0; k < cnt[i]; k++) alst2[aoff2[i] + k] = alst[aoff[i] + k];
        }
        aoff.swap(aoff2);
        alst.swap(alst2);
    }
    // chains for Rg: runs of equal molecule id in global-id order (src/main.cu:270-286)
    std::vector<int> choff;
    if (d->mol_id) {
        for (size_t i = 0; i < na; i++)
            if (i == 0 || d->mol_id[i] != d->mol_id[i - 1]) choff.push_back((int)i);
        choff.push_back((int)na);
    } else {
        choff.push_back(0);
    }
    // one device block per replica, carved into its arrays (256-byte aligned):
    // a first pass sizes the block, the second places the arrays
    int* t_boff = nullptr; int* t_blst = nullptr; int* t_aoff = nullptr; int4* t_alst = nullptr;
    int* t_choff = nullptr;
    float2* lj = nullptr; float4* bp = nullptr; float2* ap = nullptr; float* ms = nullptr;
    int* tidx = nullptr; float4* tparm = nullptr; int2* tno = nullptr; float2* tfe = nullptr;
    // tables: per-table (rmin, rmax, dr, 1/dr) with dr in FP32 as
    // src/io/table_parser.cpp:79-80, (points, offset), and (force / r, energy) rows
    const int ntab = std::max(0, d->ntables);
    std::vector<float4> h_tparm(ntab);
    std::vector<int2> h_tno(ntab);
    std::vector<float2> h_tfe;
    for (int k = 0, off = 0; k < ntab; k++) {
        const int n = d->table_n[k];
        const float lo = d->table_range[2 * k], hi = d->table_range[2 * k + 1];
        const float dr = (hi - lo) / (float)(n - 1);
        h_tparm[k] = make_float4(lo, hi, dr, 1.0f / dr);
        h_tno[k] = make_int2(n, off);
        for (int p = 0; p < n; p++)
            h_tfe.push_back(make_float2(d->table_data[3 * ((size_t)off + p) + 2],
                                        d->table_data[3 * ((size_t)off + p) + 1]));
        off += n;
    }
    char* base = nullptr;
    size_t used = 0;
    auto carve = [&](size_t bytes) -> void* {
        void* p = base ? base + used : nullptr;
        used += (std::max<size_t>(bytes, 16) + 255) & ~(size_t)255;
        return p;
    };
    auto layout = [&]() {
    D.pos = (float4*)carve(ns * 16);
    D.vel = (float4*)carve(ns * 16);
    D.force = (float4*)carve(ns * 16);
    D.pos_ref = (float4*)carve(ns * 16);
    D.ipos = (int4*)carve(ns * 16);
    D.img = (int4*)carve(ns * 16);
    D.gid = (int*)carve(ns * 4);
    D.tmp_pos = (float4*)carve(ns * 16);
    D.tmp_vel = (float4*)carve(ns * 16);
    D.tmp_img = (int4*)carve(ns * 16);
    D.cell_of = (int*)carve(ns * 4);
    D.rank = (int*)carve(ns * 4);
    D.sortx = (int*)carve(ns * 4);
    D.order = (int*)carve(ns * 4);
    D.order2 = (int*)carve(ns * 4);
    D.slot_of = (int*)carve(na * 4);
    D.cell_cnt = (int*)carve((size_t)S.ncell_cap * 4);
    D.cell_start = (int*)carve(((size_t)S.ncell_cap + 1) * 4);
    D.nnbr = (int*)carve(ns * 4);
    D.nbr = (int*)carve(nt * S.kcap * 64 * 4);
    D.gcount = (int*)carve(nt / kGroupTiles * 4);
    D.tlen = (int*)carve(nt * 4);
    D.gstage = (int*)carve(nt / kGroupTiles * kGroupStage * 4);
    D.gperm = (unsigned char*)carve(ns);
    D.bhdr = (int*)carve(ns * 4);
    D.bnd = (int*)carve(nt * S.bcap * 64 * 4);
    D.ang = (int2*)carve(nt * S.acap * 64 * 8);
    D.wpart = (double*)carve(nt * kNumPart * 8);
    D.tdisp = (float*)carve(nt * 4);
    D.nchunks = (int)((nt + kChunkTiles - 1) / kChunkTiles);
    D.cpart = (double*)carve((size_t)D.nchunks * kNumPart * 8);
    r.st = (RepState*)carve(sizeof(RepState));
    D.st = r.st;
    D.thermo = (ThermoRec*)carve((size_t)std::max(1, S.thermo_cap) * sizeof(ThermoRec));

    t_boff = (int*)carve((na + 1) * 4);
    t_blst = (int*)carve(blst.size() * 4);
    t_aoff = (int*)carve((na + 1) * 4);
    t_alst = (int4*)carve(alst.size() * 16);
    lj = (float2*)carve((size_t)d->ntypes * d->ntypes * 8);
    bp = (float4*)carve((size_t)std::max(1, d->nbond_types) * 16);
    ap = (float2*)carve((size_t)std::max(1, d->nangle_types) * 8);
    ms = (float*)carve(kMaxTypes * 4);
    t_choff = (int*)carve(choff.size() * 4);
    if (ntab) {
        tidx = (int*)carve((size_t)d->ntypes * d->ntypes * 4);
        tparm = (float4*)carve((size_t)ntab * 16);
        tno = (int2*)carve((size_t)ntab * 8);
        tfe = (float2*)carve(h_tfe.size() * 8);
    }
    D.corr = d->stress_freq > 0 ? (CorrState*)carve(sizeof(CorrState)) : nullptr;
    D.ntables = ntab;
    D.tab_idx = tidx; D.tab_parm = tparm; D.tab_no = tno; D.tab_fe = tfe;
    D.t_boff = t_boff; D.t_blst = t_blst; D.t_aoff = t_aoff; D.t_alst = t_alst;
    D.t_choff = t_choff;
    D.nchains = (int)choff.size() - 1;
    D.lj = lj; D.bparm = bp; D.aparm = ap; D.mass = ms;
    };
    layout();
    const size_t block_bytes = used;
    if (charge_bytes(block_bytes) > charge) {  // the estimate fell short of the carved layout
        charge = charge_bytes(block_bytes);
        *need = charge;
        if (in_use_ + charge > budget_) {
            delete R;
            return fail(FMD_ERR_BUDGET, "system needs " + std::to_string(charge) + " bytes; budget left " +
                                            std::to_string(budget_ - in_use_));
        }
    }
    size_t got = 0;
    base = (char*)acquire_block(block_bytes, &got);
    if (!base) {
        delete R;
        return fail(FMD_ERR_NOMEM, "device allocation failed");
    }
    r.block = base;
    r.block_bytes = got;
    r.block_used = block_bytes;
    // charged by what it needs, whether the block came fresh (whole granules)
    // or from the cache (up to 1.25x larger): the budget admits by this unit
    r.bytes = charge;
    used = 0;
    layout();


    // host staging of the per-slot state
    std::vector<float4> hpos(ns), hvel(ns);
    std::vector<int4> himg(ns);
    for (size_t i = 0; i < ns; i++) {
        if (i < na) {
            hpos[i] = make_float4(d->pos[4 * i], d->pos[4 * i + 1], d->pos[4 * i + 2], d->pos[4 * i + 3]);
            hvel[i] = d->vel ? make_float4(d->vel[3 * i], d->vel[3 * i + 1], d->vel[3 * i + 2], 0.0f)
                             : make_float4(0, 0, 0, 0);
            himg[i] = d->image ? make_int4(d->image[3 * i], d->image[3 * i + 1], d->image[3 * i + 2], (int)i)
                               : make_int4(0, 0, 0, (int)i);
        } else {
            hpos[i] = make_float4(0, 0, 0, type_bits(-1));
            hvel[i] = make_float4(0, 0, 0, 0);
            himg[i] = make_int4(0, 0, 0, 0x7fffffff);
        }
    }
    std::vector<int> ident(na);
    for (size_t i = 0; i < na; i++) ident[i] = (int)i;
    std::vector<float> mass(kMaxTypes, 1.0f);
    if (d->masses)
        for (int t = 0; t < d->ntypes; t++) mass[t] = d->masses[t];
    for (int t = 0; t < d->ntypes; t++)
        if (!(mass[t] > 0.0f)) {
            release_block(r.block, r.block_bytes);
            delete R;
            return fail(FMD_ERR_ARG, "masses must be > 0");
        }

    RepState S0;
    std::memset(&S0, 0, sizeof(S0));
    S0.step = 0;
    S0.stop = 0;
    S0.ramp_n = d->nsteps > 0 ? d->nsteps : 1;
    S0.active = 0;
    S0.rebuild = 1;
    S0.nrebuilds = 1;
    S0.L = d->L;
    S0.invL = 1.0f / d->L;
    S0.Lp = d->L;
    S0.invLp = 1.0f / d->L;
    set_grid_consts(&S0, D.rc2);
    {
        int n = (int)(d->L / D.rcs);
        int nx = n < 1 ? 1 : n;
        S0.nx = nx;
        S0.inv_cs = 1.0f / (d->L / (float)nx);
        S0.cbits = cbits_for(nx);
    }
    S0.kT = d->T_start;
    S0.vscale = 1.0f;
    S0.v_pending = 1.0f;
    S0.exp_vW = 1.0f;
    S0.fpos = d->dt;
    S0.T_start = d->T_start; S0.T_stop = d->T_stop;
    S0.P_start = d->P_start; S0.P_stop = d->P_stop;
    S0.M = std::min(std::max(1, d->chain_length), kMaxChain);
    {  // src/integrate/nose_hoover.cu:338-345, in FP64
        double Nf = 3.0 * (double)d->natoms, kT = d->T_start;
        S0.Q1 = Nf * kT * (double)d->Tdamp * (double)d->Tdamp;
        S0.Qr = kT * (double)d->Tdamp * (double)d->Tdamp;
        S0.W = (Nf + 3.0) * kT * (double)d->Pdamp * (double)d->Pdamp;
        S0.V0 = (double)d->L * (double)d->L * (double)d->L;
    }
    S0.thermo_freq = d->thermo_freq;
    S0.stress_freq = std::max(0, d->stress_freq);
    S0.thermo_cap = S.thermo_cap;

    hipStream_t s = side ? side_ : stream_;
    int32_t rc = FMD_OK;
    auto up = [&](void* dst, const void* src, size_t n) {
        if (!rc && n) rc = hip_check(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s), "upload");
    };
    auto zero = [&](void* dst, size_t n) {
        if (!rc && n) rc = hip_check(hipMemsetAsync(dst, 0, n, s), "memset");
    };
    up(D.pos, hpos.data(), ns * 16);
    up(D.pos_ref, hpos.data(), ns * 16);
    up(D.vel, hvel.data(), ns * 16);
    up(D.img, himg.data(), ns * 16);
    up(D.slot_of, ident.data(), na * 4);
    zero(D.force, ns * 16);
    zero(D.ipos, ns * 16);
    zero(D.gid, ns * 4);  // set by the sort of the initial build
    zero(D.cell_cnt, (size_t)S.ncell_cap * 4);
    zero(D.nnbr, ns * 4);
    zero(D.gcount, nt / kGroupTiles * 4);
    zero(D.tlen, nt * 4);
    zero(D.bhdr, ns * 4);
    zero(D.wpart, nt * kNumPart * 8);
    zero(D.tdisp, nt * 4);
    if (D.corr) zero(D.corr, sizeof(CorrState));
    up(t_boff, boff.data(), (na + 1) * 4);
    up(t_blst, blst.data(), blst.size() * 4);
    up(t_aoff, aoff.data(), (na + 1) * 4);
    up(t_alst, alst.data(), alst.size() * 16);
    up(lj, d->lj, (size_t)d->ntypes * d->ntypes * 8);
    if (d->nbond_types > 0) up(bp, d->bond_params, (size_t)d->nbond_types * 16);
    if (d->nangle_types > 0) up(ap, d->angle_params, (size_t)d->nangle_types * 8);
    up(ms, mass.data(), kMaxTypes * 4);
    up(t_choff, choff.data(), choff.size() * 4);
    if (ntab) {
        up(tidx, d->table_idx, (size_t)d->ntypes * d->ntypes * 4);
        up(tparm, h_tparm.data(), (size_t)ntab * 16);
        up(tno, h_tno.data(), (size_t)ntab * 8);
        up(tfe, h_tfe.data(), h_tfe.size() * 8);
    }
    up(r.st, &S0, sizeof(S0));
    up(d_descs_ + id, &D, sizeof(RepDesc));
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "sync(create)");
    if (rc) {
        release_block(r.block, r.block_bytes);
        delete R;
        return rc;
    }
    reps_[id] = R;
    in_use_ += r.bytes;
    // prime: initial build + forces (src/main.cu:311-342)
    Batch* b = batch_for({id});
    if (!b) {
        destroy(id);
        return FMD_ERR_HIP;
    }
    Launch l = launch_of(*b, s);
    launch_rebuild(l, 1, 1);
    launch_force(l, 0, b->ntl, 256, FMD_TERM_ALL, 0, FM_OBS);
    launch_boundary(l, BND_PRIME);
    rc = hip_check(hipStreamSynchronize(s), "sync(prime)");
    if (!rc) {
        RepState st;
        rc = read_state(r, st);
        if (!rc && st.err) {
            int e = st.err;
            destroy(id);
            return fail(e & ERR_NBR_OVERFLOW ? FMD_ERR_NBR_OVERFLOW : (e & ERR_STRETCH ? FMD_ERR_STRETCH : FMD_ERR_NAN),
                        "initial configuration rejected (error bits " + std::to_string(e) + ")");
        }
    }
    if (rc) {
        destroy(id);
        return rc;
    }
    *sys = id;
    *charged = charge;
    return FMD_OK;
}

int32_t Engine::destroy(int32_t sys, bool wait) {
    Replica* r = rep(sys);
    if (!r) return fail(FMD_ERR_ARG, "no such system");
    if (wait) (void)hipStreamSynchronize(stream_);
    (void)hipStreamSynchronize(side_);
    drop_batches_with(sys);
    release_block(r->block, r->block_bytes);
    in_use_ -= r->bytes;
    reps_[sys] = nullptr;
    delete r;
    return FMD_OK;
}

void Engine::drop_batches_with(int id) {
    for (auto it = batches_.begin(); it != batches_.end();) {
        if (std::find(it->first.begin(), it->first.end(), id) != it->first.end()) {
            Batch* b = it->second;
            free_batch(b);
            it = batches_.erase(it);
        } else {
            ++it;
        }
    }
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
Batch* Engine::batch_for(const std::vector<int>& ids_in) {
    std::vector<int> ids = ids_in;
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    auto it = batches_.find(ids);
    if (it != batches_.end()) return it->second;
    for (int id : ids)
        if (!rep(id)) {
            fail(FMD_ERR_ARG, "no such system " + std::to_string(id));
            return nullptr;
        }
    Batch* b = new Batch();
    b->ids = ids;
    // order replicas by size class so each class is a contiguous tile range
    std::vector<int> ord = ids;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return reps_[a]->cls < reps_[c]->cls; });
    std::vector<Tile> tiles;
    std::vector<int> t0s(ord.size());
    for (int cl = 0; cl < kNumClasses; cl++) {
        b->cls_t0[cl] = (int)tiles.size();
        for (size_t k = 0; k < ord.size(); k++)
            if (reps_[ord[k]]->cls == cl) {
                t0s[k] = (int)tiles.size();
                for (int t = 0; t < reps_[ord[k]]->ntiles; t++) tiles.push_back(Tile{ord[k], t});
            }
        b->cls_nt[cl] = (int)tiles.size() - b->cls_t0[cl];
    }
    b->ntl = (int)tiles.size();
    b->nr = (int)ord.size();
    std::vector<Tile> chunks;
    for (int id : ord)
        for (int j = 0; j < reps_[id]->desc.nchunks; j++) chunks.push_back(Tile{id, j});
    b->nch = (int)chunks.size();
    if (ord.size() > (size_t)kMaxReplicas) {
        delete b;
        fail(FMD_ERR_ARG, "more than " + std::to_string(kMaxReplicas) + " systems in one batch");
        return nullptr;
    }
    // one device block (from the replica block cache), filled by one copy
    auto al = [](size_t n) { return (std::max<size_t>(n, 16) + 255) & ~(size_t)255; };
    const size_t o_tiles = 0, o_reps = o_tiles + al(tiles.size() * sizeof(Tile)),
                 o_t0 = o_reps + al(ord.size() * 4), o_chunks = o_t0 + al(t0s.size() * 4),
                 o_nrt = o_chunks + al(chunks.size() * sizeof(Tile)), o_rlist = o_nrt + al(4),
                 o_status = o_rlist + al(tiles.size() * 4), o_phase = o_status + al(ord.size() * 16),
                 total = o_phase + al(ord.size() * 4);
    std::vector<char> host(o_rlist, 0);
    std::memcpy(host.data() + o_tiles, tiles.data(), tiles.size() * sizeof(Tile));
    std::memcpy(host.data() + o_reps, ord.data(), ord.size() * 4);
    std::memcpy(host.data() + o_t0, t0s.data(), t0s.size() * 4);
    std::memcpy(host.data() + o_chunks, chunks.data(), chunks.size() * sizeof(Tile));
    char* blk = (char*)acquire_block(total, &b->block_bytes);
    if (!blk || hipMemcpy(blk, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
        if (blk) release_block(blk, b->block_bytes);
        delete b;
        fail(FMD_ERR_NOMEM, "batch allocation failed");
        return nullptr;
    }
    b->block = blk;
    b->d_tiles = (Tile*)(blk + o_tiles);
    b->d_reps = (int*)(blk + o_reps);
    b->d_rep_t0 = (int*)(blk + o_t0);
    b->d_chunks = (Tile*)(blk + o_chunks);
    b->d_nrt = (int*)(blk + o_nrt);
    b->d_rlist = (int*)(blk + o_rlist);
    b->d_status = (int4*)(blk + o_status);
    b->d_phase = (int*)(blk + o_phase);
    b->ord = ord;
    batches_[ids] = b;
    return b;
}

// Only after the stream has drained: the block returns to the cache.
void Engine::free_batch(Batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    release_block(b->block, b->block_bytes);
    delete b;
}

Launch Engine::launch_of(const Batch& b) const { return launch_of(b, stream_); }
Launch Engine::launch_of(const Batch& b, hipStream_t s) const {
    return Launch{d_descs_, b.d_tiles, b.ntl, b.d_reps, b.nr, b.d_rep_t0, b.d_rlist, b.d_nrt, b.d_chunks, b.nch, s};
}

// Whether the batch's steps build lists beside the force pass: with enough
// replicas whose list builds are spread over enough steps of the rebuild
// period that most steps have some, read from the device (a wait for the
// stream) when the caller waits for the steps anyway; a batch in lock step
// (replicas created together rebuild together every kRebuildEvery steps)
// would pay for the cross-stream waits on every step.  The server, which
// keeps the stream busy and whose resident set changes from launch to launch
// (mostly launched without a graph), keeps the one-stream step: measured on
// its job stream, the waits cost it more than the overlap gave.
void Engine::decide_fork(Batch& b, bool may_sync) {
    b.fork_known = true;
    b.steps_since_fork = 0;
    // FMD_STEP_FORK=1 or 0 sets the structure (batches formed from then on)
    if (const char* v = std::getenv("FMD_STEP_FORK"); v && (v[0] == '0' || v[0] == '1')) {
        b.fork = v[0] == '1';
        return;
    }
    b.fork = false;
    if (b.nr < kForkReplicas || !may_sync) return;
    std::vector<int> ph(b.nr);
    launch_phase(launch_of(b), b.d_phase);
    if (hipMemcpyAsync(ph.data(), b.d_phase, ph.size() * 4, hipMemcpyDeviceToHost, stream_) != hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    bool seen[kRebuildEvery] = {};
    int distinct = 0;
    for (int v : ph)
        if (v >= 0 && !seen[v % kRebuildEvery]) {
            seen[v % kRebuildEvery] = true;
            distinct++;
        }
    b.fork = distinct >= kForkPhases;
}

// One MD step for every replica of the batch.  With many replicas, those that
// rebuild their lists this step (about one in twenty) sort their atoms, build
// their lists and compute their forces on the auxiliary stream, while the
// force pass of all the others runs on the main stream: a share of it
// (kEarlyForce) beside the sort, whose kernels are small and fit beside it,
// and the rest from the end of the sort, after the list build's workgroups
// (the larger) are queued, so they are not starved; the force pass fills the
// CUs they leave and free.  The branches touch disjoint replicas, and a
// captured graph keeps them concurrent.  A batch of few replicas gains little
// beside the cost of the cross-stream waits and runs its steps in one stream.
void Engine::launch_step_seq(Batch& b) {
    Launch l = launch_of(b);
    launch_pre(l);
    launch_decide(l);
    // the part [f0, f1) of each size class's tile range (whole groups) with
    // that class's tuned workgroup shape (adjacent classes of one shape in
    // one launch)
    auto staying = [&](int sel, double f0, double f1) {
        for (int c = 0; c < kNumClasses;) {
            int e = c + 1;
            while (e < kNumClasses && tuned_[0][e] == tuned_[0][c]) e++;
            const int t0 = b.cls_t0[c], nt = b.cls_t0[e - 1] + b.cls_nt[e - 1] - t0;
            const int a = (int)(f0 * (nt / kGroupTiles)) * kGroupTiles;
            const int z = f1 >= 1.0 ? nt : (int)(f1 * (nt / kGroupTiles)) * kGroupTiles;
            launch_force(l, t0 + a, z - a, tuned_[0][c], FMD_TERM_ALL, 1, FM_STEP, sel);
            c = e;
        }
    };
    if (b.fork) {
        const Launch la = launch_of(b, aux_);
        (void)hipEventRecord(ev_fork_, stream_);
        (void)hipStreamWaitEvent(aux_, ev_fork_, 0);
        launch_rebuild(la, 0, 0);
        (void)hipEventRecord(ev_sorted_, aux_);
        launch_build(la, kForkBuildBlocks);
        launch_force_listed(la, 1, FM_STEP);
        (void)hipEventRecord(ev_join_, aux_);
        staying(FS_STAYING, 0.0, kEarlyForce);
        (void)hipStreamWaitEvent(stream_, ev_sorted_, 0);
        staying(FS_STAYING, kEarlyForce, 1.0);
        (void)hipStreamWaitEvent(stream_, ev_join_, 0);
    } else {
        launch_rebuild(l, 0, 1);
        staying(FS_ALL, 0.0, 1.0);
    }
    launch_boundary(l, BND_STEP);
}

int32_t Engine::begin_run(const std::vector<int>& ids, int64_t nsteps) {
    Batch* b = batch_for(ids);
    if (!b) return FMD_ERR_ARG;
    Launch l = launch_of(*b);
    launch_control(l, CTL_RUN, nsteps);
    launch_boundary(l, BND_START);
    return hip_check(hipGetLastError(), "launch(begin_run)");
}

int32_t Engine::begin_run_side(const std::vector<int>& ids, int64_t nsteps) {
    Batch* b = batch_for(ids);
    if (!b) return FMD_ERR_ARG;
    Launch l = launch_of(*b, side_);
    launch_control(l, CTL_RUN, nsteps);
    launch_boundary(l, BND_START);
    int32_t rc = hip_check(hipGetLastError(), "launch(begin_run)");
    if (!rc) rc = hip_check(hipStreamSynchronize(side_), "sync(side)");
    return rc;
}

int32_t Engine::launch_steps(const std::vector<int>& ids, int64_t k, bool use_graph, bool may_sync) {
    Batch* b = batch_for(ids);
    if (!b) return FMD_ERR_ARG;
    // FMD_NO_GRAPH=1 issues the same kernel sequence without graph capture
    // (for tracing tools that do not follow graph launches)
    static const bool no_graph = [] {
        const char* v = std::getenv("FMD_NO_GRAPH");
        return v && v[0] == '1';
    }();
    if (no_graph) use_graph = false;
    // the step structure: chosen with the batch, again with each capture, and
    // again by a waiting caller every kForkRecheck steps (the replicas' list
    // builds may have spread out or fallen into step since)
    if (!b->fork_known || (may_sync && b->steps_since_fork >= kForkRecheck)) {
        const bool was = b->fork;
        decide_fork(*b, may_sync);
        if (b->graph && b->fork != was) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    }
    b->steps_since_fork += k;
    if (use_graph) {
        if (!b->graph || b->tune_version != tune_version_) {
            if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
            decide_fork(*b, may_sync);
            hipGraph_t g;
            int32_t rc = hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "BeginCapture");
            if (rc) return rc;
            for (int s = 0; s < kGraphSteps; s++) launch_step_seq(*b);
            rc = hip_check(hipStreamEndCapture(stream_, &g), "EndCapture");
            if (rc) return rc;
            rc = hip_check(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0), "GraphInstantiate");
            (void)hipGraphDestroy(g);
            if (rc) return rc;
            // the graph's one-time setup now, not in its first launch
            rc = hip_check(hipGraphUpload(b->graph, stream_), "GraphUpload");
            if (rc) return rc;
            b->graph_steps = kGraphSteps;
            b->tune_version = tune_version_;
        }
        int64_t full = k / b->graph_steps;
        for (int64_t g = 0; g < full; g++) {
            int32_t rc = hip_check(hipGraphLaunch(b->graph, stream_), "GraphLaunch");
            if (rc) return rc;
        }
        k -= full * b->graph_steps;
    }
    for (int64_t s = 0; s < k; s++) launch_step_seq(*b);
    return hip_check(hipGetLastError(), "launch(step)");
}

int32_t Engine::sync() { return hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

int32_t Engine::read_state(Replica& r, RepState& s) {
    return hip_check(hipMemcpy(&s, r.st, sizeof(RepState), hipMemcpyDeviceToHost), "read state");
}

static int32_t err_code(int e) {
    if (e & ERR_NAN) return FMD_ERR_NAN;
    if (e & ERR_NBR_OVERFLOW) return FMD_ERR_NBR_OVERFLOW;
    if (e & ERR_STRETCH) return FMD_ERR_STRETCH;
    if (e & ERR_CELL) return FMD_ERR_CELL;
    if (e & ERR_BONDED) return FMD_ERR_ARG;
    if (e & ERR_INTERNAL) return FMD_ERR_STATE;
    return FMD_OK;
}

int32_t Engine::check_errors(const std::vector<int>& ids, std::vector<int>* failed) {
    // one status kernel and one copy for the whole batch
    Batch* b = batch_for(ids);
    if (!b) return FMD_ERR_ARG;
    std::vector<int4> stat(b->nr);
    launch_status(launch_of(*b), b->d_status);
    int32_t rc = hip_check(hipMemcpyAsync(stat.data(), b->d_status, stat.size() * sizeof(int4),
                                          hipMemcpyDeviceToHost, stream_), "copy status");
    if (!rc) rc = sync();
    if (rc) return rc;
    int32_t first = FMD_OK;
    for (int k = 0; k < b->nr; k++) {
        const int id = b->ord[k];
        Replica* r = rep(id);
        const int4 st = stat[k];
        r->steps_done = (int64_t)(((uint64_t)(uint32_t)st.w << 32) | (uint32_t)st.z);
        if (st.x) {
            if (failed) failed->push_back(id);
            if (!first) {
                first = err_code(st.x);
                err_ = "system " + std::to_string(id) + " stopped at step " + std::to_string(r->steps_done) +
                       " (error bits " + std::to_string(st.x) + ")";
            }
        }
    }
    return first;
}

int32_t Engine::replica_error(int32_t sys) {
    Replica* r = rep(sys);
    if (!r) return FMD_ERR_ARG;
    RepState s;
    if (int32_t rc = read_state(*r, s)) return rc;
    return err_code(s.err);
}

void Engine::trim_batches(const std::vector<int>& keep_in) {
    std::vector<int> keep = keep_in;
    std::sort(keep.begin(), keep.end());
    keep.erase(std::unique(keep.begin(), keep.end()), keep.end());
    for (auto it = batches_.begin(); it != batches_.end();) {
        if (it->first != keep) {
            Batch* b = it->second;
            free_batch(b);
            it = batches_.erase(it);
        } else {
            ++it;
        }
    }
}

int32_t Engine::step(const int32_t* ids, int32_t n, int64_t nsteps, bool use_graph) {
    if (n <= 0 || !ids) return fail(FMD_ERR_ARG, "empty system list");
    if (nsteps < 0) return fail(FMD_ERR_ARG, "negative step count");
    std::vector<int> v(ids, ids + n);
    for (int id : v)
        if (!rep(id)) return fail(FMD_ERR_ARG, "no such system " + std::to_string(id));
    if (nsteps == 0) return FMD_OK;
    if (int32_t rc = begin_run(v, nsteps)) return rc;
    if (int32_t rc = launch_steps(v, nsteps, use_graph, true)) return rc;
    if (int32_t rc = sync()) return rc;
    return check_errors(v, nullptr);
}

int64_t Engine::step_count(int32_t sys) {
    Replica* r = rep(sys);
    if (!r) return -1;
    RepState s;
    if (read_state(*r, s)) return -1;
    return s.step;
}
int64_t Engine::rebuilds(int32_t sys) {
    Replica* r = rep(sys);
    if (!r) return -1;
    RepState s;
    if (read_state(*r, s)) return -1;
    return s.nrebuilds;
}

int32_t Engine::set_force_cap(int32_t sys, float cap) {
    Replica* r = rep(sys);
    if (!r) return fail(FMD_ERR_ARG, "no such system");
    if (int32_t rc = sync()) return rc;
    r->desc.fcap = cap > 0.0f ? cap : 0.0f;
    r->desc.vlim = r->desc.fcap > 0.0f ? 0.02f / r->desc.dt : 0.0f;
    return hip_check(hipMemcpy(d_descs_ + sys, &r->desc, sizeof(RepDesc), hipMemcpyHostToDevice), "upload desc");
}

int32_t Engine::set_langevin(int32_t sys, float T0, float T1, float Tdamp) {
    Replica* r = rep(sys);
    if (!r || !(Tdamp > 0.0f)) return fail(FMD_ERR_ARG, "bad arguments");
    if (r->ensemble != FMD_LANGEVIN) return fail(FMD_ERR_STATE, "system is not Langevin");
    if (int32_t rc = sync()) return rc;
    r->desc.c1 = std::exp(-(1.0f / Tdamp) * r->desc.dt);
    r->desc.c2 = std::sqrt(1.0f - r->desc.c1 * r->desc.c1);
    RepState s;
    if (int32_t rc = read_state(*r, s)) return rc;
    s.T_start = T0;
    s.T_stop = T1;
    s.pre_ready = 0;  // recompute the current step's target
    if (int32_t rc = hip_check(hipMemcpy(r->st, &s, sizeof(RepState), hipMemcpyHostToDevice), "upload state")) return rc;
    return hip_check(hipMemcpy(d_descs_ + sys, &r->desc, sizeof(RepDesc), hipMemcpyHostToDevice), "upload desc");
}

// ---------------------------------------------------------------------------
// parity entry points
// ---------------------------------------------------------------------------
int32_t Engine::build_neighbors(int32_t sys) {
    Replica* r = rep(sys);
    if (!r) return fail(FMD_ERR_ARG, "no such system");
    Batch* b = batch_for({sys});
    if (!b) return FMD_ERR_ARG;
    Launch l = launch_of(*b);
    launch_control(l, CTL_FORCE_REBUILD, 0);
    launch_rebuild(l, 1, 1);
    // slots were permuted: recompute the forces in the new order
    launch_force(l, 0, b->ntl, 256, FMD_TERM_ALL, 0, FM_OBS);
    launch_boundary(l, BND_REDUCE);
    if (int32_t rc = sync()) return rc;
    RepState s;
    if (int32_t rc = read_state(*r, s)) return rc;
    if (s.err) return fail(err_code(s.err), "neighbor build failed (error bits " + std::to_string(s.err) + ")");
    return FMD_OK;
}

int32_t Engine::get_pairs(int32_t sys, int64_t* npairs, int32_t* ij, int64_t cap) {
    Replica* r = rep(sys);
    if (!r || !npairs) return fail(FMD_ERR_ARG, "bad arguments");
    const RepDesc& D = r->desc;
    std::vector<int> nbr((size_t)D.ntiles * D.kcap * 64), nn(D.nslots);
    std::vector<int4> img(D.nslots);
    int32_t rc = hip_check(hipMemcpy(nbr.data(), D.nbr, nbr.size() * 4, hipMemcpyDeviceToHost), "copy nbr");
    if (!rc) rc = hip_check(hipMemcpy(nn.data(), D.nnbr, nn.size() * 4, hipMemcpyDeviceToHost), "copy nnbr");
    if (!rc) rc = hip_check(hipMemcpy(img.data(), D.img, img.size() * 16, hipMemcpyDeviceToHost), "copy img");
    if (rc) return rc;
    // entries are LDS byte offsets into the group's neighbor set (gcount >= 0) or slots
    const int ng = D.ntiles / kGroupTiles;
    std::vector<int> gcount(ng), gstage((size_t)ng * kGroupStage);
    rc = hip_check(hipMemcpy(gcount.data(), D.gcount, gcount.size() * 4, hipMemcpyDeviceToHost), "copy gcount");
    if (!rc) rc = hip_check(hipMemcpy(gstage.data(), D.gstage, gstage.size() * 4, hipMemcpyDeviceToHost), "copy gstage");
    if (rc) return rc;
    std::vector<unsigned char> perm(D.nslots);
    rc = hip_check(hipMemcpy(perm.data(), D.gperm, perm.size(), hipMemcpyDeviceToHost), "copy gperm");
    if (rc) return rc;
    std::vector<std::pair<int, int>> pairs;
    for (int col = 0; col < D.nslots; col++) {
        // the row in column col belongs to atom i of the column's group
        const int lt = col / 64, lane = col % 64, g = lt / kGroupTiles;
        const int i = g * kGroupSlots + perm[col];
        if (i >= D.natoms) continue;
        const uint16_t* row16 = (const uint16_t*)(nbr.data() + (size_t)lt * D.kcap * 64);
        const uint16_t* gs16 = (const uint16_t*)gstage.data() + (size_t)g * kGroupStage;  // as on the device
        for (int k = 0; k < nn[i]; k++) {
            int j = D.idx16 ? (int)row16[(size_t)k * 64 + lane] : nbr[((size_t)lt * D.kcap + k) * 64 + lane];
            if (gcount[g] >= 0) {
                const int q = j >> 4;
                if ((j & 15) || q >= gcount[g]) return fail(FMD_ERR_STATE, "corrupt neighbor entry");
                j = D.idx16 ? (int)gs16[q] : gstage[(size_t)g * kGroupStage + q];
            }
            if (j < 0 || j >= D.natoms) return fail(FMD_ERR_STATE, "corrupt neighbor entry");
            if (j == i) continue;  // a candidate the list build rejected after writing it
            pairs.emplace_back(img[i].w, img[j].w);
        }
    }
    std::sort(pairs.begin(), pairs.end());
    *npairs = (int64_t)pairs.size();
    if (ij)
        for (int64_t p = 0; p < (int64_t)pairs.size() && p < cap; p++) {
            ij[2 * p] = pairs[p].first;
            ij[2 * p + 1] = pairs[p].second;
        }
    return FMD_OK;
}

int32_t Engine::get_cell_ids(int32_t sys, int32_t* out, int32_t* nx) {
    Replica* r = rep(sys);
    if (!r || !out) return fail(FMD_ERR_ARG, "bad arguments");
    int* dout = nullptr;
    int32_t rc = hip_check(hipMalloc(&dout, (size_t)r->natoms * 4), "hipMalloc");
    if (rc) return rc;
    launch_ref_cells(d_descs_ + sys, r->natoms, dout, stream_);
    rc = sync();
    if (!rc) rc = hip_check(hipMemcpy(out, dout, (size_t)r->natoms * 4, hipMemcpyDeviceToHost), "copy");
    (void)hipFree(dout);
    if (!rc && nx) {
        RepState s;
        rc = read_state(*r, s);
        int n = (int)(s.L / r->desc.rcs);
        *nx = n < 1 ? 1 : n;
    }
    return rc;
}

int32_t Engine::get_morton(int32_t sys, uint32_t* out) {
    Replica* r = rep(sys);
    if (!r || !out) return fail(FMD_ERR_ARG, "bad arguments");
    uint32_t* dout = nullptr;
    int32_t rc = hip_check(hipMalloc(&dout, (size_t)r->natoms * 4), "hipMalloc");
    if (rc) return rc;
    launch_morton_keys(d_descs_ + sys, r->natoms, dout, stream_);
    rc = sync();
    if (!rc) rc = hip_check(hipMemcpy(out, dout, (size_t)r->natoms * 4, hipMemcpyDeviceToHost), "copy");
    (void)hipFree(dout);
    return rc;
}

int32_t Engine::compute_forces(int32_t sys, uint32_t terms, float* f4, double* vir, double* pe) {
    Replica* r = rep(sys);
    if (!r) return fail(FMD_ERR_ARG, "no such system");
    Batch* b = batch_for({sys});
    if (!b) return FMD_ERR_ARG;
    Launch l = launch_of(*b);
    launch_force(l, 0, b->ntl, 256, (int)(terms & 7u), 0, FM_OBS);
    launch_boundary(l, BND_REDUCE);
    int32_t rc = sync();
    RepState s;
    if (!rc) rc = read_state(*r, s);
    if (!rc && f4) {
        const RepDesc& D = r->desc;
        std::vector<float4> f(D.nslots);
        std::vector<int4> img(D.nslots);
        rc = hip_check(hipMemcpy(f.data(), D.force, f.size() * 16, hipMemcpyDeviceToHost), "copy force");
        if (!rc) rc = hip_check(hipMemcpy(img.data(), D.img, img.size() * 16, hipMemcpyDeviceToHost), "copy img");
        if (!rc)
            for (int i = 0; i < D.natoms; i++) {
                int g = img[i].w;
                f4[4 * g] = f[i].x; f4[4 * g + 1] = f[i].y; f4[4 * g + 2] = f[i].z; f4[4 * g + 3] = f[i].w;
            }
    }
    if (!rc) {
        if (vir)
            for (int c = 0; c < 6; c++) vir[c] = s.vir[c];
        if (pe) *pe = s.pe;
    }
    if ((terms & 7u) != 7u) {  // on every path: restore the full forces the next step's kick uses
        launch_force(l, 0, b->ntl, 256, FMD_TERM_ALL, 0, FM_OBS);
        launch_boundary(l, BND_REDUCE);
        const int32_t rc2 = sync();
        if (!rc) rc = rc2;
    }
    if (!rc && (s.err & ERR_STRETCH)) return fail(FMD_ERR_STRETCH, "bond stretched beyond R0");
    return rc;
}

// ---------------------------------------------------------------------------
// observation
// ---------------------------------------------------------------------------
int32_t Engine::get_state(int32_t sys, float* pos4, float* vel3, int32_t* img3, bool wait) {
    Replica* r = rep(sys);
    if (!r) return fail(FMD_ERR_ARG, "no such system");
    const RepDesc& D = r->desc;
    std::vector<float4> p(D.nslots), v(D.nslots);
    std::vector<int4> img(D.nslots);
    RepState s;
    int32_t rc = wait ? sync() : FMD_OK;
    if (!rc) rc = hip_check(hipMemcpy(p.data(), D.pos, p.size() * 16, hipMemcpyDeviceToHost), "copy pos");
    if (!rc) rc = hip_check(hipMemcpy(v.data(), D.vel, v.size() * 16, hipMemcpyDeviceToHost), "copy vel");
    if (!rc) rc = hip_check(hipMemcpy(img.data(), D.img, img.size() * 16, hipMemcpyDeviceToHost), "copy img");
    if (!rc) rc = read_state(*r, s);
    if (rc) return rc;
    const float vs = s.v_pending;  // post-force scale not yet applied to vel[]
    for (int i = 0; i < D.natoms; i++) {
        int g = img[i].w;
        if (pos4) { pos4[4 * g] = p[i].x; pos4[4 * g + 1] = p[i].y; pos4[4 * g + 2] = p[i].z; pos4[4 * g + 3] = p[i].w; }
        if (vel3) { vel3[3 * g] = v[i].x * vs; vel3[3 * g + 1] = v[i].y * vs; vel3[3 * g + 2] = v[i].z * vs; }
        if (img3) { img3[3 * g] = img[i].x; img3[3 * g + 1] = img[i].y; img3[3 * g + 2] = img[i].z; }
    }
    return FMD_OK;
}

int32_t Engine::get_box(int32_t sys, double* L, bool of_positions) {
    Replica* r = rep(sys);
    if (!r || !L) return fail(FMD_ERR_ARG, "bad arguments");
    RepState s;
    int32_t rc = read_state(*r, s);
    if (!rc) *L = of_positions ? s.Lp : s.L;
    return rc;
}

// Ensemble average of the chains' Rg (src/analysis/rg.cu:106-129), summed on the
// host in chain order.
int32_t Engine::get_rg(int32_t sys, double* rg, double* per_chain, int32_t cap, int32_t* nchains) {
    Replica* r = rep(sys);
    if (!r || !rg) return fail(FMD_ERR_ARG, "bad arguments");
    const int nc = r->desc.nchains;
    if (nchains) *nchains = nc;
    *rg = 0.0;
    if (nc <= 0) return FMD_OK;
    double* dout = nullptr;
    int32_t rc = hip_check(hipMalloc(&dout, (size_t)nc * 8), "hipMalloc(rg)");
    if (rc) return rc;
    launch_rg(d_descs_ + sys, nc, dout, stream_);
    std::vector<double> h(nc);
    rc = hip_check(hipMemcpyAsync(h.data(), dout, (size_t)nc * 8, hipMemcpyDeviceToHost, stream_), "copy rg");
    if (!rc) rc = sync();
    (void)hipFree(dout);
    if (rc) return rc;
    double s = 0.0;
    for (int c = 0; c < nc; c++) s += h[c];
    *rg = s / nc;
    if (per_chain)
        for (int c = 0; c < nc && c < cap; c++) per_chain[c] = h[c];
    return FMD_OK;
}

int32_t Engine::get_rg_noalloc(int32_t sys, double* rg, double* per_chain, int32_t cap, int32_t* nchains) {
    Replica* r = rep(sys);
    if (!r || !rg) return fail(FMD_ERR_ARG, "bad arguments");
    const int nc = r->desc.nchains;
    if (nchains) *nchains = nc;
    *rg = 0.0;
    if (nc <= 0) return FMD_OK;
    if (!scratch_) return fail(FMD_ERR_HIP, "the device scratch array has no address on this device");
    if (nc > kRgScratchChains)
        return fail(FMD_ERR_ARG, "system has " + std::to_string(nc) + " chains; the allocation-free Rg read holds " +
                                     std::to_string(kRgScratchChains));
    launch_rg(d_descs_ + sys, nc, scratch_, stream_);
    launch_rg_mean(d_descs_ + sys, scratch_, scratch_ + kRgScratchChains, stream_);
    int32_t rc = hip_check(hipGetLastError(), "launch(rg)");
    if (!rc) rc = hip_check(hipMemcpyAsync(rg, scratch_ + kRgScratchChains, 8, hipMemcpyDeviceToHost, stream_), "copy rg");
    if (!rc && per_chain && cap > 0)
        rc = hip_check(hipMemcpyAsync(per_chain, scratch_, (size_t)std::min(nc, cap) * 8, hipMemcpyDeviceToHost, stream_),
                       "copy rg");
    if (!rc) rc = sync();
    return rc;
}

// ---------------------------------------------------------------------------
// Snapshot and resume.  A replica's whole state lives in its device block
// (per-slot arrays in the current sorted order, the Verlet list and its
// reference positions, the thermostat / barostat and step counters, the
// correlator and the thermo ring, the topology and parameters), and its
// descriptor points into that block only.  A snapshot is a header, the
// descriptor with every pointer stored as an offset into the block, and the
// block's bytes; resuming uploads the block into a new one and rebases the
// pointers, so the resumed replica continues bit for bit.
// ---------------------------------------------------------------------------
namespace {
constexpr uint32_t kSnapMagic = 0x534D4446u;  // "FDMS"
constexpr uint32_t kSnapVersion = 3;
struct SnapHeader {
    uint32_t magic, version, desc_bytes, state_bytes;
    uint64_t used;
    int32_t natoms, nslots, ntiles, cls, ensemble, thermo_cap;
    int64_t steps_done, thermo_read;
    RepDesc desc;  // pointers as (offset into the block + 1), 0 = null
};

// Applies f to every pointer member of a descriptor.
template <class F>
void for_each_ptr(RepDesc& d, F&& f) {
    f(d.pos); f(d.vel); f(d.force); f(d.pos_ref); f(d.ipos); f(d.img); f(d.gid);
    f(d.tmp_pos); f(d.tmp_vel); f(d.tmp_img); f(d.cell_of); f(d.rank); f(d.sortx); f(d.order); f(d.order2);
    f(d.slot_of); f(d.cell_cnt); f(d.cell_start); f(d.nnbr); f(d.nbr); f(d.gcount); f(d.tlen); f(d.gstage); f(d.gperm); f(d.bhdr); f(d.bnd); f(d.ang);
    f(d.t_boff); f(d.t_blst); f(d.t_aoff); f(d.t_alst); f(d.t_choff);
    f(d.lj); f(d.bparm); f(d.aparm); f(d.mass);
    f(d.tab_idx); f(d.tab_parm); f(d.tab_no); f(d.tab_fe); f(d.corr);
    f(d.wpart); f(d.tdisp); f(d.cpart); f(d.st); f(d.thermo);
}
}  // namespace

int32_t Engine::snapshot_size(int32_t sys, uint64_t* bytes) {
    Replica* r = rep(sys);
    if (!r || !bytes) return fail(FMD_ERR_ARG, "bad arguments");
    *bytes = sizeof(SnapHeader) + r->block_used;
    return FMD_OK;
}

int32_t Engine::snapshot(int32_t sys, void* buf, uint64_t cap) {
    Replica* r = rep(sys);
    if (!r || !buf) return fail(FMD_ERR_ARG, "bad arguments");
    if (cap < sizeof(SnapHeader) + r->block_used) return fail(FMD_ERR_ARG, "snapshot buffer too small");
    SnapHeader h;
    std::memset(&h, 0, sizeof(h));
    h.magic = kSnapMagic; h.version = kSnapVersion;
    h.desc_bytes = sizeof(RepDesc); h.state_bytes = sizeof(RepState);
    h.used = r->block_used;
    h.natoms = r->natoms; h.nslots = r->nslots; h.ntiles = r->ntiles; h.cls = r->cls;
    h.ensemble = r->ensemble; h.thermo_cap = r->thermo_cap;
    h.steps_done = r->steps_done; h.thermo_read = r->thermo_read;
    h.desc = r->desc;
    const char* base = (const char*)r->block;
    bool outside = false;
    for_each_ptr(h.desc, [&](auto*& p) {
        const char* q = (const char*)p;
        if (!q) return;
        if (q < base || q >= base + r->block_used) { outside = true; return; }
        p = reinterpret_cast<std::remove_reference_t<decltype(p)>>((uintptr_t)(q - base) + 1);
    });
    if (outside) return fail(FMD_ERR_ARG, "internal: descriptor points outside the replica block");
    int32_t rc = sync();
    if (!rc) rc = hip_check(hipMemcpy((char*)buf + sizeof(SnapHeader), r->block, r->block_used,
                                      hipMemcpyDeviceToHost), "copy snapshot");
    if (rc) return rc;
    std::memcpy(buf, &h, sizeof(h));
    return FMD_OK;
}

int32_t Engine::resume(const void* buf, uint64_t size, int32_t* sys) {
    if (!buf || !sys || size < sizeof(SnapHeader)) return fail(FMD_ERR_ARG, "bad arguments");
    SnapHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (h.magic != kSnapMagic || h.version != kSnapVersion || h.desc_bytes != sizeof(RepDesc) ||
        h.state_bytes != sizeof(RepState))
        return fail(FMD_ERR_ARG, "not a snapshot of this library version");
    if (size != sizeof(SnapHeader) + h.used) return fail(FMD_ERR_ARG, "snapshot size does not match its header");
    bool bad = false;
    for_each_ptr(h.desc, [&](auto*& p) {
        const uintptr_t o = (uintptr_t)p;
        if (o > h.used) bad = true;
    });
    if (bad || h.natoms <= 0 || h.nslots < h.natoms || h.cls < 0 || h.cls >= kNumClasses || !h.desc.st)
        return fail(FMD_ERR_ARG, "corrupt snapshot header");
    if (in_use_ + charge_bytes(h.used) > budget_)
        return fail(FMD_ERR_BUDGET, "snapshot needs " + std::to_string(h.used) + " bytes; budget left " +
                                        std::to_string(budget_ - in_use_));
    int id = -1;
    for (int k = 0; k < kMaxReplicas; k++)
        if (!reps_[k]) { id = k; break; }
    if (id < 0) return fail(FMD_ERR_BUDGET, "replica table full");
    size_t got = 0;
    char* base = (char*)acquire_block(h.used, &got);
    if (!base) return fail(FMD_ERR_NOMEM, "device allocation failed");
    Replica* R = new Replica();
    Replica& r = *R;
    r.id = id;
    r.natoms = h.natoms; r.nslots = h.nslots; r.ntiles = h.ntiles; r.cls = h.cls;
    r.ensemble = h.ensemble; r.thermo_cap = h.thermo_cap;
    r.steps_done = h.steps_done; r.thermo_read = h.thermo_read;
    r.block = base; r.block_bytes = got; r.block_used = h.used; r.bytes = charge_bytes(h.used);
    r.desc = h.desc;
    for_each_ptr(r.desc, [&](auto*& p) {
        const uintptr_t o = (uintptr_t)p;
        p = o ? reinterpret_cast<std::remove_reference_t<decltype(p)>>(base + (o - 1)) : nullptr;
    });
    r.st = r.desc.st;
    int32_t rc = hip_check(hipMemcpyAsync(base, (const char*)buf + sizeof(SnapHeader), h.used,
                                          hipMemcpyHostToDevice, stream_), "upload snapshot");
    if (!rc) rc = hip_check(hipMemcpyAsync(d_descs_ + id, &r.desc, sizeof(RepDesc), hipMemcpyHostToDevice, stream_),
                            "upload descriptor");
    if (!rc) rc = sync();
    if (rc) {
        release_block(r.block, r.block_bytes);
        delete R;
        return rc;
    }
    reps_[id] = R;
    in_use_ += r.bytes;
    *sys = id;
    return FMD_OK;
}

// Stress autocorrelator (src/analysis/correlator.cu:24-32,94-110).
int32_t Engine::corr_push(int32_t sys, const double* x) {
    Replica* r = rep(sys);
    if (!r || !x) return fail(FMD_ERR_ARG, "bad arguments");
    if (!r->desc.corr) return fail(FMD_ERR_ARG, "system has no stress correlator (stress_freq = 0)");
    Sample6 v;
    for (int c = 0; c < 6; c++) v.v[c] = x[c];
    launch_corr_push(d_descs_ + sys, v, stream_);
    return hip_check(hipGetLastError(), "corr push");
}

int32_t Engine::corr_reset(int32_t sys) {
    Replica* r = rep(sys);
    if (!r) return fail(FMD_ERR_ARG, "bad arguments");
    if (!r->desc.corr) return fail(FMD_ERR_ARG, "system has no stress correlator (stress_freq = 0)");
    return hip_check(hipMemsetAsync(r->desc.corr, 0, sizeof(CorrState), stream_), "corr reset");
}

int32_t Engine::get_corr(int32_t sys, double* corr, int32_t* counts, int32_t* levels) {
    Replica* r = rep(sys);
    if (!r) return fail(FMD_ERR_ARG, "bad arguments");
    if (!r->desc.corr) return fail(FMD_ERR_ARG, "system has no stress correlator (stress_freq = 0)");
    CorrState h;
    int32_t rc = hip_check(hipMemcpyAsync(&h, r->desc.corr, sizeof(CorrState), hipMemcpyDeviceToHost, stream_),
                           "copy corr");
    if (!rc) rc = sync();
    if (rc) return rc;
    if (corr) std::memcpy(corr, h.corr, sizeof(h.corr));
    if (counts) std::memcpy(counts, h.counts, sizeof(h.counts));
    if (levels) *levels = kCorrLevels;
    return FMD_OK;
}

static void thermo_of(const RepState& s, int natoms, fmd_thermo* o) {
    const double N = natoms, L = s.L, V = L * L * L;
    o->step = s.rec_step;
    o->ke = 0.5 * (s.kin[0] + s.kin[3] + s.kin[5]) / N;
    o->pe = s.pe / N;
    o->temp = 2.0 * o->ke / 3.0;
    for (int c = 0; c < 6; c++) o->press[c] = (s.kin[c] + s.vir[c]) / V;
    o->L = L;
    float m2;
    int g;
    max_key_split(s.vkey, &m2, &g);
    o->max_v = std::sqrt((double)m2) * (double)s.v_pending;
    o->max_v_id = g;
    max_key_split(s.fkey, &m2, &g);
    o->max_f = std::sqrt((double)m2);
    o->max_f_id = g;
}

int32_t Engine::get_thermo(int32_t sys, fmd_thermo* out) {
    Replica* r = rep(sys);
    if (!r || !out) return fail(FMD_ERR_ARG, "bad arguments");
    RepState s;
    int32_t rc = read_state(*r, s);
    if (!rc && s.obs_step != s.step) {
        // PE and virial are evaluated only on observed steps; evaluate them now
        Batch* b = batch_for({sys});
        if (!b) return FMD_ERR_ARG;
        Launch l = launch_of(*b);
        launch_force(l, 0, b->ntl, 256, FMD_TERM_ALL, 0, FM_OBS);
        launch_boundary(l, BND_REDUCE);
        rc = sync();
        if (!rc) rc = read_state(*r, s);
    }
    if (!rc) thermo_of(s, r->natoms, out);
    return rc;
}

int32_t Engine::get_thermostat(int32_t sys, fmd_thermostat* out) {
    Replica* r = rep(sys);
    if (!r || !out) return fail(FMD_ERR_ARG, "bad arguments");
    RepState s;
    int32_t rc = read_state(*r, s);
    if (rc) return rc;
    *out = fmd_thermostat{};
    out->ensemble = r->ensemble;
    out->M = s.M;
    out->step = s.step;
    out->Q1 = s.Q1;
    out->Qr = s.Qr;
    out->W = s.W;
    out->V0 = s.V0;
    out->eps = s.eps;
    out->veps = s.veps;
    for (int k = 0; k < kMaxChain && k < FMD_MAX_CHAIN; k++) {
        out->xi[k] = s.xi[k];
        out->vxi[k] = s.vxi[k];
    }
    out->ke_carry = s.ke_carry;
    return FMD_OK;
}

int32_t Engine::nh_chain_run(const fmd_chain_run* c, double* xi, double* vxi, double* scales) {
    if (!c || !xi || !vxi || !scales || (c->n > 0 && !c->ke) || c->n < 0 || c->M < 1 ||
        c->M > FMD_MAX_CHAIN || c->natoms < 1)
        return fail(FMD_ERR_ARG, "bad chain-run arguments");
    if (c->n == 0) return FMD_OK;
    const size_t nsc = (size_t)c->n * (c->carry ? 2 : 1);
    RepState h{};
    h.M = c->M;
    h.Q1 = c->Q1;
    h.Qr = c->Qr;
    for (int k = 0; k < kMaxChain; k++) {
        h.xi[k] = xi[k];
        h.vxi[k] = vxi[k];
    }
    void* buf = nullptr;
    const size_t bytes = sizeof(RepState) + (size_t)c->n * 8 + nsc * 8;
    int32_t rc = hip_check(hipMalloc(&buf, bytes), "hipMalloc(chain run)");
    if (rc) return rc;
    RepState* dst = (RepState*)buf;
    double* dke = (double*)((char*)buf + sizeof(RepState));
    double* dsc = dke + c->n;
    rc = hip_check(hipMemcpyAsync(dst, &h, sizeof(RepState), hipMemcpyHostToDevice, stream_), "copy chain");
    if (!rc) rc = hip_check(hipMemcpyAsync(dke, c->ke, (size_t)c->n * 8, hipMemcpyHostToDevice, stream_), "copy ke");
    if (!rc) {
        launch_nh_chain_run(dst, dke, c->n, c->carry ? 1 : 0, c->kT, c->natoms, c->half_dt, dsc, stream_);
        rc = hip_check(hipGetLastError(), "chain run launch");
    }
    if (!rc) rc = hip_check(hipMemcpyAsync(scales, dsc, nsc * 8, hipMemcpyDeviceToHost, stream_), "copy scales");
    if (!rc) rc = hip_check(hipMemcpyAsync(&h, dst, sizeof(RepState), hipMemcpyDeviceToHost, stream_), "copy chain");
    if (!rc) rc = sync();
    (void)hipFree(buf);
    if (rc) return rc;
    for (int k = 0; k < kMaxChain; k++) {
        xi[k] = h.xi[k];
        vxi[k] = h.vxi[k];
    }
    return FMD_OK;
}

int32_t Engine::drain_thermo(int32_t sys, fmd_thermo* out, int32_t cap, int32_t* n, int64_t* lost) {
    Replica* r = rep(sys);
    if (!r || !n) return fail(FMD_ERR_ARG, "bad arguments");
    *n = 0;
    if (lost) *lost = 0;
    if (r->thermo_cap == 0) return FMD_OK;
    RepState s;
    int32_t rc = read_state(*r, s);
    if (rc) return rc;
    int64_t from = r->thermo_read, to = s.thermo_written;
    if (to - from > r->thermo_cap) {
        if (lost) *lost = to - from - r->thermo_cap;
        from = to - r->thermo_cap;
    }
    if (to - from > cap) to = from + cap;
    if (to <= from) return FMD_OK;
    std::vector<ThermoRec> ring(r->thermo_cap);
    rc = hip_check(hipMemcpy(ring.data(), r->desc.thermo, ring.size() * sizeof(ThermoRec), hipMemcpyDeviceToHost), "copy thermo");
    if (rc) return rc;
    for (int64_t k = from; k < to; k++) {
        const ThermoRec& t = ring[k % r->thermo_cap];
        fmd_thermo& o = out[*n];
        o.step = t.step; o.ke = t.ke; o.pe = t.pe; o.temp = t.temp;
        for (int c = 0; c < 6; c++) o.press[c] = t.press[c];
        o.L = t.L;
        o.max_v = t.maxv; o.max_f = t.maxf; o.max_v_id = t.maxv_id; o.max_f_id = t.maxf_id;
        (*n)++;
    }
    r->thermo_read = to;
    return FMD_OK;
}

int32_t Engine::time_force(const int32_t* ids, int32_t n, uint32_t terms, int reps, double* ms,
                           double* bytes, int64_t* pairs) {
    if (!ids || n <= 0 || reps < 1) return fail(FMD_ERR_ARG, "bad arguments");
    std::vector<int> v(ids, ids + n);
    Batch* b = batch_for(v);
    if (!b) return FMD_ERR_ARG;
    Launch l = launch_of(*b);
    // algorithmic bytes from the current lists and bonded tables
    double B = 0.0;
    int64_t P = 0;
    for (int id : b->ids) {
        Replica* r = rep(id);
        std::vector<int> nn(r->nslots), hdr(r->nslots);
        int32_t rc = hip_check(hipMemcpy(nn.data(), r->desc.nnbr, nn.size() * 4, hipMemcpyDeviceToHost), "copy");
        if (!rc) rc = hip_check(hipMemcpy(hdr.data(), r->desc.bhdr, hdr.size() * 4, hipMemcpyDeviceToHost), "copy");
        if (rc) return rc;
        int64_t p = 0, nb = 0, na = 0;
        for (int i = 0; i < r->natoms; i++) { p += nn[i]; nb += hdr[i] & 0xFF; na += (hdr[i] >> 8) & 0xFF; }
        P += p;
        // per bead: 16 B position, 4 B row count, 16 B force; per listed pair:
        // a 4-byte index, whatever width the rows are stored in (the position
        // gathers are cache traffic)
        if (terms & 1) B += 36.0 * r->natoms + 4.0 * (double)p;
        if (terms & 6) B += 4.0 * r->natoms + ((terms & 1) ? 0.0 : 32.0 * r->natoms);
        if (terms & 2) B += 4.0 * (double)nb;
        if (terms & 4) B += 8.0 * (double)na;
    }
    hipEvent_t e0, e1;
    if (int32_t rc = hip_check(hipEventCreate(&e0), "event")) return rc;
    if (int32_t rc = hip_check(hipEventCreate(&e1), "event")) return rc;
    // one launch over every tile (block shape of the class holding most tiles),
    // so a profiler's per-dispatch average is the figure returned here
    int big = 0;
    for (int c = 1; c < kNumClasses; c++)
        if (b->cls_nt[c] > b->cls_nt[big]) big = c;
    const int blk = tuned_[0][big];
    auto run = [&]() { launch_force(l, 0, b->ntl, blk, (int)(terms & 7u), 0, FM_FAST); };
    run();
    (void)hipEventRecord(e0, stream_);
    for (int k = 0; k < reps; k++) run();
    (void)hipEventRecord(e1, stream_);
    int32_t rc = hip_check(hipEventSynchronize(e1), "event sync");
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    // restore the full forces the next step's kick reads
    launch_force(l, 0, b->ntl, 256, FMD_TERM_ALL, 0, FM_OBS);
    if ((rc = sync())) return rc;
    if (ms) *ms = (double)t / reps;
    if (bytes) *bytes = B;
    if (pairs) *pairs = P;
    return FMD_OK;
}

// ---------------------------------------------------------------------------
// autotuner: per size class, time the force kernel over the class's tiles for
// each candidate block shape and keep the fastest (a workgroup of 1, 2 or 4
// waves lies inside one group; smaller ones stage the group's set more often).
// The list build has one shape: one workgroup per group of tiles.
// ---------------------------------------------------------------------------
// Per size class: the force pass's neighbor layout (staged neighbor sets or
// direct gathers; the list build writes the layout, so each is timed after a
// rebuild in it) and its workgroup shape.  The candidates are timed on copies
// of the systems (snapshot and resume), so the systems themselves keep their
// lists and rebuild schedule; every candidate gives the same bits.  The cost
// of a layout is its force pass plus a tenth of its list build (a system
// rebuilds about every ten steps).  A new layout reaches each system at its
// next rebuild.
int32_t Engine::autotune(const int32_t* ids, int32_t n, int reps) {
    if (!ids || n <= 0) return fail(FMD_ERR_ARG, "empty system list");
    if (reps < 1) reps = 1;
    std::vector<int> tmp;
    int32_t rc = sync();
    for (int k = 0; k < n && !rc; k++) {
        uint64_t sz = 0;
        rc = snapshot_size(ids[k], &sz);
        std::vector<char> buf(sz);
        if (!rc) rc = snapshot(ids[k], buf.data(), sz);
        int32_t t = -1;
        if (!rc) rc = resume(buf.data(), sz, &t);
        if (!rc) tmp.push_back(t);
    }
    Batch* b = rc ? nullptr : batch_for(tmp);
    if (!rc && !b) rc = FMD_ERR_ARG;
    const int cand[3] = {64, 128, 256};  // the shapes the kernels are compiled for (__launch_bounds__(256))
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) rc = hip_check(hipEventCreate(&e0), "event");
    if (!rc) rc = hip_check(hipEventCreate(&e1), "event");
    auto timed = [&](auto&& run) -> double {
        run();  // warm
        (void)hipEventRecord(e0, stream_);
        for (int k = 0; k < reps; k++) run();
        (void)hipEventRecord(e1, stream_);
        if (!rc) rc = hip_check(hipGetLastError(), "autotune launch");
        if (!rc) rc = hip_check(hipEventSynchronize(e1), "autotune sync");
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return (double)ms / reps;
    };
    int best_layout[kNumClasses], best_block[kNumClasses];
    double best_ms[kNumClasses];
    for (int c = 0; c < kNumClasses; c++) {
        best_layout[c] = tuned_[2][c];
        best_block[c] = tuned_[0][c];
        best_ms[c] = 1e300;
        if (b && b->cls_nt[c] > 0) tune_rows_[c].clear();
    }
    for (int layout = 1; layout >= 0 && !rc; layout--) {
        for (int t : tmp) {
            Replica* r = rep(t);
            r->desc.nostage = layout ? 0 : 1;
            rc = hip_check(hipMemcpy(d_descs_ + t, &r->desc, sizeof(RepDesc), hipMemcpyHostToDevice), "desc");
            if (rc) break;
        }
        if (rc) break;
        Launch l = launch_of(*b);
        launch_rebuild(l, 1, 1);
        rc = sync();
        // the list build of the whole batch in this layout, shared out by tiles
        const double build_ms = rc ? 0.0 : timed([&] { launch_rebuild(l, 1, 1); });
        for (int c = 0; c < kNumClasses && !rc; c++) {
            if (b->cls_nt[c] == 0) continue;
            const double share = build_ms * (double)b->cls_nt[c] / (double)b->ntl;
            for (int bi = 0; bi < 3 && !rc; bi++) {
                const double ms = timed([&] {
                    launch_force(l, b->cls_t0[c], b->cls_nt[c], cand[bi], FMD_TERM_ALL, 0, FM_FAST);
                });
                tune_rows_[c].push_back(TuneRow{layout, cand[bi], ms});
                const double cost = ms + 0.1 * share;
                if (cost < best_ms[c]) {
                    best_ms[c] = cost;
                    best_layout[c] = layout;
                    best_block[c] = cand[bi];
                }
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (int t : tmp) destroy(t);
    if (rc) return rc;
    for (int c = 0; c < kNumClasses; c++) {
        if (!b->cls_nt[c]) continue;
        tuned_[0][c] = best_block[c];
        if (tuned_[2][c] != best_layout[c]) {
            tuned_[2][c] = best_layout[c];
            apply_layout(c);
        }
    }
    tune_version_++;
    return FMD_OK;
}

void Engine::apply_layout(int cls) {
    for (Replica* r : reps_)
        if (r && r->cls == cls) {
            r->desc.nostage = tuned_[2][cls] ? 0 : 1;
            (void)hipMemcpy(d_descs_ + r->id, &r->desc, sizeof(RepDesc), hipMemcpyHostToDevice);
        }
}

int32_t Engine::tune_times(int cls, int32_t* layout, int32_t* block, double* ms, int32_t cap) const {
    if (cls < 0 || cls >= kNumClasses) return -1;
    const auto& rows = tune_rows_[cls];
    for (int k = 0; k < (int)rows.size() && k < cap; k++) {
        if (layout) layout[k] = rows[k].layout;
        if (block) block[k] = rows[k].block;
        if (ms) ms[k] = rows[k].ms;
    }
    return (int32_t)rows.size();
}

int32_t Engine::set_tuned(int kernel, int cls, int value) {
    if (kernel < 0 || kernel > 2 || cls < 0 || cls >= kNumClasses)
        return fail(FMD_ERR_ARG, "kernel must be 0, 1 or 2 and the size class in [0, 3)");
    if (kernel == 0 && value != 64 && value != 128 && value != 256)
        return fail(FMD_ERR_ARG, "block shape must be 64, 128 or 256");
    if (kernel == 1 && value != kGroupSlots)
        return fail(FMD_ERR_ARG, "the list build runs one workgroup of " + std::to_string(kGroupSlots) +
                                     " threads per group of tiles");
    if (kernel == 2 && value != 0 && value != 1)
        return fail(FMD_ERR_ARG, "the neighbor layout is 1 (staged) or 0 (direct)");
    if (int32_t rc = sync()) return rc;
    tuned_[kernel][cls] = value;
    if (kernel == 2) apply_layout(cls);
    tune_version_++;
    return FMD_OK;
}

}  // namespace fmd
