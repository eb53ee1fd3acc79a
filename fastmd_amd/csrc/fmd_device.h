// fmd_device.h -- data layout shared by the host engine and the gfx950 kernels.
//
// Every replica (one simulation system) owns its own device buffers.  A launch
// covers many replicas through two small device tables:
//   RepDesc[r] : pointers and constants of replica r (written at creation)
//   Tile[t]    : (replica, tile) for each 64-slot wave tile in the launch
// A tile never straddles replicas (slot counts are padded to 64), so every
// per-wave reduction belongs to exactly one replica and results do not depend
// on which other replicas share the launch.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fmd {

constexpr int kTile = 64;           // slots per tile = one wave64
constexpr int kMaxReplicas = 4096;  // entries of an engine's replica table
// Groups of kGroupTiles consecutive tiles of a replica (a replica's slot count
// is a multiple of the group size) share one neighborhood.  The list build
// stages a group's candidate atoms -- the union of the 27-neighborhoods of the
// cells holding its atoms, at most kMaxUnion cells and kBuildStage atoms -- in
// LDS.  The atoms the group's rows actually reference (its neighbor set, at
// most kGroupStage) are listed per group; the force kernel stages them in LDS
// and the rows hold their LDS byte offsets.
constexpr int kGroupTiles = 4;
constexpr int kGroupSlots = kGroupTiles * kTile;
constexpr int kMaxUnion = 256;
constexpr int kBuildStage = 3072;
constexpr int kGroupStage = 1696;
constexpr int kMaxTypes = 16;
constexpr int kMaxChain = 10;
constexpr int kNumPart = 16;   // FP64 partials per tile: pe, W[6], K[6], |v| and |F| max keys, pad
constexpr int kNumSum = 13;    // the first kNumSum partials are sums, the next two maxima
constexpr int kChunkTiles = 256;  // tiles summed by one workgroup in the first reduction level
constexpr int kRebuildEvery = 20;  // forced rebuild period (src/main.cu:401)
constexpr int kSlotBits = 26;      // slot / gid field of packed bonded entries
constexpr int kSlotMask = (1 << kSlotBits) - 1;

enum ErrBits : int {
    ERR_NBR_OVERFLOW = 1,
    ERR_STRETCH = 2,
    ERR_NAN = 4,
    ERR_CELL = 8,
    ERR_BONDED = 16,
    ERR_INTERNAL = 32,  // an internal consistency check failed
};

struct Tile {
    int rep;   // replica index into the RepDesc table
    int lt;    // tile index within the replica
};

// Thermo record written by the device every thermo_freq steps.
struct ThermoRec {
    int64_t step;
    double ke, pe, temp;
    double press[6];
    double L;
    double maxv, maxf;         // largest |v| and |F| of one atom
    int64_t maxv_id, maxf_id;  // its global id (the smallest on ties)
};

// Largest-magnitude key of a wave or replica reduction: the FP32 |x|^2 in the
// high word and the complement of the global id in the low word, so the
// largest key is the largest magnitude with the smallest id; read as a
// double it is finite and positive and orders like the integer (the
// reductions take fmax of doubles).
__host__ __device__ inline double max_key(float mag2, int gid) {
    union { float f; uint32_t u; } m;
    m.f = mag2;
    union { uint64_t u; double d; } k;
    k.u = ((uint64_t)m.u << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)gid);
    return k.d;
}
__host__ __device__ inline void max_key_split(double key, float* mag2, int* gid) {
    union { uint64_t u; double d; } k;
    k.d = key;
    union { float f; uint32_t u; } m;
    m.u = (uint32_t)(k.u >> 32);
    *mag2 = m.f;
    *gid = (k.u == 0) ? -1 : (int)(0xFFFFFFFFu - (uint32_t)k.u);
}

// Mutable per-replica scalar state.  Written by per-replica kernels (one lane),
// read by the per-atom kernels of the next phase.
struct RepState {
    // step control
    int64_t step;          // steps completed
    int64_t stop;          // replica runs while step < stop
    int64_t ramp_n;        // denominator of the T/P ramp (run.conf nsteps)
    int active;            // 1 while stepping
    int rebuild;           // rebuild decision of the current step
    int since_rebuild;
    int maxdisp_bits;      // max |r - r_ref|^2 since the last rebuild (float bits)
    int err;               // ErrBits
    int pre_ready;         // start-of-step scalars for step+1 are computed
    int want_obs;          // the current step records PE and virial (NPT or thermo step)
    int kicked;            // Langevin: the last force pass also gave the next step's
                           // pre-force half kick (vel holds it; force was not stored)
    int64_t nrebuilds;
    // current box and cell grid (float copies for per-atom kernels)
    float L, invL;      // box of the next position update (the barostat advances it after the forces)
    float Lp, invLp;    // box the current positions were placed in
    float glo, giu2;    // LJ grid-unit constants of Lp (set_grid_consts)
    unsigned gspan;
    int nx;                // cells per axis
    int cbits;             // Morton bits per axis of the cell key
    float inv_cs;
    // per-step factors for the pre-force kernel
    float kT;              // Langevin target of the current step
    float vscale;          // velocity scale owed before the kick (NH / NPT)
    float exp_vW, fpos;    // NPT drift: r = r*exp_vW + v*fpos, pos_ref *= exp_vW
    float v_pending;       // post-force scale already owed (applied on readback)
    // FP64 thermostat / barostat state
    double T_start, T_stop, P_start, P_stop;
    double xi[kMaxChain], vxi[kMaxChain];
    int M;
    double Q1, Qr, W, V0, eps, veps;
    double ke_carry;
    // last reductions (after the post-force kick and scaling)
    double pe, vir[6], kin[6];
    double vkey, fkey;     // max_key of |v|^2 (before the pending scale) and |F|^2
    int64_t rec_step;      // step of the last reduction
    int64_t obs_step;      // step of the last PE / virial evaluation
    // thermo ring
    int thermo_freq;
    int thermo_cap;
    int stress_freq;       // correlator sample interval (0 = off)
    int64_t thermo_written;  // total records written
};

// Multiple-tau correlator of one replica (src/analysis/correlator.cu:33-92),
// FP64 sums: per level a ring of kCorrPoints samples, the lag sums and counts,
// and the block accumulator feeding the next level every kCorrPoints samples.
constexpr int kCorrLevels = 16, kCorrPoints = 16;
struct CorrState {
    double buf[kCorrLevels * kCorrPoints * 6];
    double corr[kCorrLevels * kCorrPoints * 6];
    double acc[kCorrLevels * 6];
    int counts[kCorrLevels * kCorrPoints];
    int ins[kCorrLevels];
    int bcount[kCorrLevels];
};

struct RepDesc {
    int natoms, nslots, ntiles, ntypes;
    int kcap;              // neighbor rows per tile
    int bcap, acap;        // bonded rows per tile
    int ncell_cap;         // capacity of the Morton cell-key arrays
    int ensemble;
    float rc2, rcs, rcs2, half_skin2;
    float dt, half_dt;
    float c1, c2;          // Langevin: exp(-dt/Tdamp), sqrt(1-c1^2)
    float fcap;            // LJ pair force cap (0 = off)
    float vlim;            // push-off speed limit (0 = off): |v| dt <= 0.02 sigma
    int clamp_lj;          // force/r clamp to +-1000 (src/force/lj.cu:48)
    float lj_kf, lj_fclamp;  // one (eps, sigma): 24 eps / sigma^2 and 1000 / that (0 otherwise)
    int idx16;             // neighbor rows hold 16-bit slots (nslots <= 65536)
    int nostage;           // the list build leaves rows as slots (the force kernel gathers
                           // neighbors from HBM) instead of offsets into a staged neighbor set
    uint32_t seed_lo, seed_hi;
    // per-slot state (slot = position in the cell-sorted order)
    float4* pos;
    float4* vel;
    float4* force;
    float4* pos_ref;
    int4* ipos;            // fixed-point copy of pos for the force path (x, y, z, type bits)
    int4* img;             // image counts x, y, z and the global id in w
    int* gid;              // the global id alone (img.w), for the per-step noise key
    float4* tmp_pos;
    float4* tmp_vel;
    int4* tmp_img;
    int* cell_of;
    int* rank;
    int* sortx;            // [nslots] x grid coordinate relative to the atom's cell (the within-cell sort key)
    int* order;
    int* order2;
    int* slot_of;          // [natoms] global id -> slot
    int* cell_cnt;         // [ncell_cap]
    int* cell_start;       // [ncell_cap + 1]
    int* nnbr;             // [nslots]
    int* nbr;              // [ntiles][kcap][64] entries: 16-bit if idx16, else 32-bit; the LDS byte
                           // offset of the neighbor in its group's staged set, or its slot
                           // when the group's gcount is -1
    int* gcount;           // [ntiles / kGroupTiles] atoms of the group's neighbor set, -1: rows hold slots
    int* tlen;             // [ntiles] padded length of every row of the tile (its longest)
    int* gstage;           // [ntiles / kGroupTiles][kGroupStage] slots of the neighbor set (16-bit if idx16)
    unsigned char* gperm;  // [nslots] atom (index in its group) whose row is in column (tile, lane);
                           // a group's rows are placed by descending length
    int* bhdr;             // [nslots] nbonds | nangles << 8
    int* bnd;              // [ntiles][bcap][64] partner slot | type << 26
    int2* ang;             // [ntiles][acap][64] (slot | role << 26, slot | type << 26)
    // topology by global id (static)
    const int* t_boff;     // [natoms + 1]
    const int* t_blst;     // partner gid | type << 26
    const int* t_aoff;     // [natoms + 1]
    const int4* t_alst;    // angle (i, j, k, type) incident on the gid
    const int* t_choff;    // [nchains + 1] chains: runs of equal molecule id in global-id order
    int nchains;
    // parameters
    const float2* lj;      // [ntypes^2] (eps, sig)
    const float4* bparm;   // per bond type (k, R0, eps, sig)
    const float2* aparm;   // per angle type (k, theta0)
    const float* mass;     // per type
    // tabulated pair potentials (src/force/table.cu:6-182): a type pair has an
    // LJ entry or a table; rows are interpolated linearly in r
    int ntables;
    const int* tab_idx;    // [ntypes^2] table of the pair, -1 = none
    const float4* tab_parm;  // per table: rmin, rmax, dr, 1/dr
    const int2* tab_no;    // per table: points, offset into tab_fe
    const float2* tab_fe;  // per point: force / r, energy
    CorrState* corr;       // stress autocorrelator (null when off)
    // reductions and records
    double* wpart;         // [ntiles][kNumPart]
    float* tdisp;          // [ntiles] largest squared displacement since the last rebuild (-1: non-finite)
    double* cpart;         // [nchunks][kNumPart] partials summed over chunks of kChunkTiles tiles
    int nchunks;
    RepState* st;
    ThermoRec* thermo;     // [thermo_cap]
};

// ---------------------------------------------------------------------------
// Geometry.  The pair-selection arithmetic is evaluated without contraction so
// it is bit-identical to the reference formulas on the CPU.
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

__host__ __device__ inline int type_of(float w) { return __builtin_bit_cast(int, w); }
__host__ __device__ inline float type_bits(int t) { return __builtin_bit_cast(float, t); }

// min_image of src/core/pbc.cuh:4-6
// dr - L * rint(dr * invL) (src/core/pbc.cuh:4-12).  For wrapped positions
// |dr| <= L, so the image shift m = rint(dr * invL) is -1, 0 or 1 and L * m is
// exact: the single-rounding fma(-L, m, dr) returns the same bits as the
// two-rounding form, in one instruction fewer.
__device__ inline float min_image_ref(float dr, float L, float invL) {
    return fmaf(-L, rintf(dr * invL), dr);
}
__device__ inline float dist2_ref(float4 a, float4 b, float L, float invL) {
    float dx = min_image_ref(a.x - b.x, L, invL);
    float dy = min_image_ref(a.y - b.y, L, invL);
    float dz = min_image_ref(a.z - b.z, L, invL);
    return dx * dx + dy * dy + dz * dz;
}
// Force-path coordinates: each axis of a position is kept as an unsigned
// 32-bit fixed-point number on a grid of 2^32 steps over the box,
// X = floor(x * 2^32 / L) mod 2^32 (FP64 product).  The difference of two such
// numbers in two's complement is the minimum-image displacement, exact on the
// grid; it is converted to FP32 and scaled by L (one rounding each).  The grid
// step L / 2^32 is 1e-8 sigma for a 40-sigma box.
__device__ inline unsigned fix_coord(float x, double scale) {
    return (unsigned)(long long)floor((double)x * scale);
}
// ulen = L * 2^-32 is exact (power-of-two scaling), so (float)d * ulen rounds
// once, identically to ((float)d * L) * 2^-32.
__device__ inline float fix_diff(unsigned a, unsigned b, float ulen) {
    return (float)(int)(a - b) * ulen;
}

// LJ cutoff test in grid units.  a = (float)(int)(Xi - Xj) per axis (the
// fixed-point displacement in grid steps u = L * 2^-32, rounded once) and
// s = fma(az, az, fma(ay, ay, fma(ax, ax, -lo_g))), which is
// (r^2 - 1e-10) / u^2 up to rounding.  A pair is inside when
// 0 < s < span, with lo_g = (float)(1e-10f / u^2) and
// span = (float)((rc^2 - 1e-10f) / u^2), both quotients in FP64 (u^2 is exact
// in FP64): the r^2 < rc^2 && r^2 > 1e-10 test of src/force/lj.cu:37 up to
// one FP32 rounding of each side.  s stands in for r^2 / u^2 in the force
// terms (the two differ by a relative 1e-10 / r^2).  As unsigned integers the
// bit patterns of non-negative floats order as the floats and a negative s
// (sign bit set) lies above every bound, so 0 < s < span is one subtract and
// one unsigned compare, and 0 <= s < span a single compare.
struct GridCut { float lo_g; unsigned span; };
// The grid-unit constants of a box of length Lp, kept in the replica's state
// and set wherever Lp changes (creation, each step's rebuild decision), so
// the force pass reads them instead of dividing: lo_g, span (as its bit
// pattern) and 1/u^2, quotients in FP64 (the same bits on host and device).
__host__ __device__ inline void set_grid_consts(RepState* st, float rc2) {
    const float u = st->Lp * 2.3283064365386963e-10f;  // L * 2^-32, exact
    const double u2 = (double)u * (double)u;
    const float span = (float)(((double)rc2 - (double)1e-10f) / u2);
    st->glo = (float)((double)1e-10f / u2);
    __builtin_memcpy(&st->gspan, &span, sizeof(float));
    st->giu2 = (float)(1.0 / u2);
}
__device__ inline GridCut grid_cut(const RepState* st) { return GridCut{st->glo, st->gspan}; }
__device__ inline float grid_d(unsigned a, unsigned b) { return (float)(int)(a - b); }
__device__ inline float grid_s(float x, float y, float z, float lo_g) {
    return fmaf(z, z, fmaf(y, y, fmaf(x, x, -lo_g)));
}
__device__ inline bool grid_inside(float s, unsigned span) { return __float_as_uint(s) - 1u < span - 1u; }
// 0 <= s < span: also takes s == 0 (r^2 = 1e-10 to rounding) as inside, for
// the forces-only LJ loop, whose clamped force/r keeps that pair finite.
__device__ inline bool grid_inside0(float s, unsigned span) { return __float_as_uint(s) < span; }

// Morton interleave of 10 bits (src/core/morton.cu:4-10)
__host__ __device__ inline uint32_t expand_bits(uint32_t v) {
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__host__ __device__ inline uint32_t cell_key(int cx, int cy, int cz) {
    return (expand_bits((uint32_t)cz) << 2) | (expand_bits((uint32_t)cy) << 1) |
           expand_bits((uint32_t)cx);
}
// Cell coordinate of one axis (src/neighbor/verlet_list.cu:67-72)
__device__ inline int cell_coord(float x, float inv_cs, int nx) {
    int c = (int)floorf(x * inv_cs);
    return ((c % nx) + nx) % nx;
}

#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and Box-Muller; bit-identical to the
// host restatement for the integer part.
// ---------------------------------------------------------------------------
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += W0; k1 += W1; }
        uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
        uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}
__device__ inline float u01(uint32_t x) { return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f); }
__device__ inline float3 normal3(uint32_t seed_lo, uint32_t seed_hi, uint32_t gid, int64_t step) {
    uint32_t c[4] = {gid, (uint32_t)step, (uint32_t)((uint64_t)step >> 32), 0x4C414E47u};
    philox4x32_10(c, seed_lo, seed_hi);
    const float twopi = 6.28318530717958647692f;
    float a = sqrtf(-2.0f * logf(u01(c[0])));
    float b = sqrtf(-2.0f * logf(u01(c[2])));
    float s1, c1;
    sincosf(twopi * u01(c[1]), &s1, &c1);
    return make_float3(a * c1, a * s1, b * cosf(twopi * u01(c[3])));
}

}  // namespace fmd
