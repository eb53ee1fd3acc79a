// kernels.hip -- gfx950 kernels of the bead-spring MD step.
//
// Per-atom kernels run one wave64 per 64-slot tile.  Per-replica kernels run
// one workgroup (or lane) per replica.  Every kernel reads the replica's
// RepState to decide whether it has work (active / rebuild flags), so a fixed
// launch sequence can be captured once into a hipGraph and replayed while
// replicas rebuild, stop or fail independently.
//
// Step sequence (src/main.cu:403-630, one launch each, all replicas at once):
//   pre       kick + drift (+ Langevin O step, NH/NPT scaling), wrap, skin max
//   decide    per replica: rebuild if max displacement > (skin/2)^2 or every 20
//   bin/scan/scatter/rank/gather   Morton-ordered cell sort (rebuilding only)
//   build     Verlet list + bonded slot tables (rebuilding only)
//   force     LJ + FENE/WCA + angle, post-force kick, FP64 wave partials
//   boundary  per replica: FP64 reductions, NH chain / barostat, thermo record
#include "kernels.h"

#include <climits>

namespace fmd {

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------
__device__ inline int wave_in_grid() {
    return __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
}
__device__ inline int waves_in_grid() { return gridDim.x * (blockDim.x >> 6); }
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Renumbering block b to (b % 8) * (G / 8) + b / 8 (for the first G - G % 8
// blocks) gives each XCD one contiguous range of tiles, so a replica's
// positions are gathered through one L2 instead of eight.
constexpr int kXcds = 8;
__device__ inline int xcd_block() {
    const int b = blockIdx.x, g = gridDim.x, g8 = g - g % kXcds;
    return b < g8 ? (b % kXcds) * (g8 / kXcds) + b / kXcds : b;
}
__device__ inline int xcd_wave_in_grid() {
    return __builtin_amdgcn_readfirstlane(xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6));
}
__device__ inline int lane_id() { return threadIdx.x & 63; }
// set bits of m below this lane
__device__ inline int lane_prefix(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ inline float wave_max(float x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ inline double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
// The sum over the wave in lane 63 (other lanes hold partial sums), by DPP
// moves on the VALU instead of LDS permutes: pairs, quads (quad_perm), 8 and
// 16 lanes (row_shr 4, 8), then rows 0+1 and 2+3 (row_bcast 15) and all four
// (row_bcast 31).  Every lane of the wave must be active.
template <int CTRL, int ROWS>
__device__ inline double dpp_move(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWS, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWS, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double wave_sum_last(double x) {
    x += dpp_move<0xB1, 0xF>(x);   // quad_perm [1, 0, 3, 2]
    x += dpp_move<0x4E, 0xF>(x);   // quad_perm [2, 3, 0, 1]
    x += dpp_move<0x114, 0xF>(x);  // row_shr 4
    x += dpp_move<0x118, 0xF>(x);  // row_shr 8
    x += dpp_move<0x142, 0xA>(x);  // row_bcast 15 into rows 1 and 3
    x += dpp_move<0x143, 0xC>(x);  // row_bcast 31 into rows 2 and 3
    return x;
}
// The largest value of the wave in lane 63 (non-negative inputs: the lanes a
// DPP move leaves without a source read 0).
__device__ inline double wave_max_last(double x) {
    x = fmax(x, dpp_move<0xB1, 0xF>(x));
    x = fmax(x, dpp_move<0x4E, 0xF>(x));
    x = fmax(x, dpp_move<0x114, 0xF>(x));
    x = fmax(x, dpp_move<0x118, 0xF>(x));
    x = fmax(x, dpp_move<0x142, 0xA>(x));
    x = fmax(x, dpp_move<0x143, 0xC>(x));
    return x;
}
__device__ inline int wave_min_i(int x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ inline int wave_max_i(int x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ inline int wave_or(int x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
    return x;
}

// Global address-space views of descriptor pointers, so the hot loops issue
// global_load instead of flat_load.
#if defined(__HIP_DEVICE_COMPILE__)
#define FMD_GLOBAL __attribute__((address_space(1)))
#else
#define FMD_GLOBAL
#endif
template <class T> __device__ inline const FMD_GLOBAL T* gview(const T* p) {
    return (const FMD_GLOBAL T*)p;
}
// The same for stores.  A flat store also counts against the LDS wait
// counter, so a flat store ahead of LDS reads makes their wait include the
// store's trip to memory; a global store does not.
template <class T> __device__ inline FMD_GLOBAL T* gmut(T* p) { return (FMD_GLOBAL T*)p; }

constexpr int kMaxBlockK = 256;  // launch bound of k_build and k_force
constexpr int kLJChunk = 8;      // neighbors per software-pipelined chunk (kcap % 8 == 0)
constexpr int kLJTail = 4;       // rows are padded to multiples of this; a last chunk may be this short
constexpr int kBuildChunk = 8;   // list-build candidates whose loads are issued together
constexpr int kBndThreads = 256;  // workgroup of the per-replica boundary reduction

#pragma clang fp contract(off)
__device__ inline float sq3(float x, float y, float z) { return x * x + y * y + z * z; }
#pragma clang fp contract(fast)

// Word 3 of a raw buffer resource on gfx9 parts: 32-bit data format, no swizzle.
constexpr int kBufferWord3 = 0x00020000;

// Indexed ("structured") buffer load of one 16-byte record: the resource
// carries the stride, the hardware forms index * stride (buffer_load_dwordx4
// ... idxen), and an index at or past the record count reads zeros.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
__device__ u32x4 struct_load_b128(__amdgpu_buffer_rsrc_t rsrc, unsigned vindex, unsigned voffset,
                                  unsigned soffset, int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.v4i32");

// ---------------------------------------------------------------------------
// FP64 Nose-Hoover chain (src/integrate/nose_hoover.cu:86-178 / :356-447) and
// NPT barostat (src/integrate/nose_hoover.cu:185-305), one lane per replica.
// ---------------------------------------------------------------------------
__device__ double nh_chain(RepState* st, double KE, double kT, int natoms, double half_dt) {
    const int M = st->M;
    double xi[kMaxChain], v[kMaxChain], dd[kMaxChain];
    for (int k = 0; k < kMaxChain; k++) { xi[k] = st->xi[k]; v[k] = st->vxi[k]; dd[k] = 0.0; }
    const double Nf = 3.0 * (double)natoms, Q1 = st->Q1, Qr = st->Qr;
    const double dt2 = half_dt, dt4 = 0.5 * half_dt, dt8 = 0.25 * half_dt;
    double T = 2.0 * KE / Nf;
    if (Q1 > 0.0) dd[0] = (Nf * T - Nf * kT) / Q1;
    for (int ich = M - 1; ich > 0; ich--) {
        double e = (ich + 1 < M) ? exp(-dt8 * v[ich + 1]) : 1.0;
        v[ich] = (v[ich] * e + dd[ich] * dt4) * e;
    }
    double e0 = (M > 1) ? exp(-dt8 * v[1]) : 1.0;
    v[0] = (v[0] * e0 + dd[0] * dt4) * e0;
    double scale = exp(-dt2 * v[0]);
    T *= scale * scale;
    if (Q1 > 0.0) dd[0] = (Nf * T - Nf * kT) / Q1;
    for (int k = 0; k < M; k++) xi[k] += dt2 * v[k];
    v[0] = (v[0] * e0 + dd[0] * dt4) * e0;
    for (int ich = 1; ich < M; ich++) {
        double e = (ich + 1 < M) ? exp(-dt8 * v[ich + 1]) : 1.0;
        double Qp = (ich == 1) ? Q1 : Qr;
        dd[ich] = (Qp * v[ich - 1] * v[ich - 1] - kT) / Qr;
        v[ich] = (v[ich] * e + dd[ich] * dt4) * e;
    }
    for (int k = 0; k < kMaxChain; k++) { st->xi[k] = xi[k]; st->vxi[k] = v[k]; }
    return scale;
}

struct BaroOut {
    double L, baro_scale, exp_vW, veps_W_dt, nh_scale, KE_scaled;
};
__device__ BaroOut npt_chain_baro(RepState* st, double KE, double vir_trace, bool advance,
                                  double kT, double Pt, int natoms, double half_dt, double dt) {
    BaroOut o;
    const double Nf = 3.0 * (double)natoms;
    double V = st->V0 * exp(3.0 * st->eps);
    if (advance) {
        double P = (2.0 * KE + vir_trace) / (3.0 * V);
        double cb = cbrt(2.0);
        double w[3] = {1.0 / (2.0 - cb), -cb / (2.0 - cb), 1.0 / (2.0 - cb)};
        for (int s = 0; s < 3; s++) {
            double ws = w[s] * dt;
            double dv = 3.0 * V * (P - Pt) * ws;
            double vh = st->veps + 0.5 * dv;
            st->eps += vh * ws / st->W;
            dv = 3.0 * st->V0 * exp(3.0 * st->eps) * (P - Pt) * ws;
            st->veps += dv;
        }
        V = st->V0 * exp(3.0 * st->eps);
    }
    o.L = cbrt(V);
    double vW = st->veps / st->W;
    o.baro_scale = exp(-(1.0 + 3.0 / Nf) * vW * half_dt);
    o.exp_vW = exp(vW * dt);
    o.veps_W_dt = vW * dt;
    o.KE_scaled = KE * o.baro_scale * o.baro_scale;
    o.nh_scale = nh_chain(st, o.KE_scaled, kT, natoms, half_dt);
    return o;
}

__device__ inline double sinch(double x) {
    if (fabs(x) < 1e-6) return 1.0 + x * x / 6.0;
    return sinh(x) / x;
}

// Cell grid of the current box (src/neighbor/verlet_list.cu:8-13), recomputed
// at every rebuild so the grid follows the box under NPT.
__device__ void set_grid(const RepDesc& d, RepState* st) {
    int n = (int)(st->Lp / d.rcs);
    int nx = n < 1 ? 1 : n;
    float cs = st->Lp / (float)nx;
    int cb = 0;
    while ((1 << cb) < nx) cb++;
    if ((1 << (3 * cb)) > d.ncell_cap) {
        st->err |= ERR_CELL;
        st->active = 0;
        return;
    }
    st->nx = nx;
    st->inv_cs = 1.0f / cs;
    st->cbits = cb;
}

// Start-of-step scalars for step s (T/P ramp, NH chain, barostat factors).
__device__ void start_step(const RepDesc& d, RepState* st, int64_t s) {
    float frac = (float)(s - 1) / (float)st->ramp_n;  // src/main.cu:407-409
    float T = (float)st->T_start + ((float)st->T_stop - (float)st->T_start) * frac;
    float P = (float)st->P_start + ((float)st->P_stop - (float)st->P_start) * frac;
    const double hdt = 0.5 * (double)d.dt, dt = (double)d.dt;
    // observables on NPT steps (the barostat needs the virial), thermo steps and
    // the last step of a run (its result is read back)
    st->want_obs = (d.ensemble == 2) || (st->thermo_freq > 0 && (s % st->thermo_freq) == 0) ||
                   (st->stress_freq > 0 && (s % st->stress_freq) == 0) || (s == st->stop);
    if (d.ensemble == 0) {
        st->kT = T;
        st->vscale = 1.0f;
    } else if (d.ensemble == 1) {
        double nh = nh_chain(st, st->ke_carry, (double)T, d.natoms, hdt);
        st->vscale = (float)((double)st->v_pending * nh);
    } else {
        BaroOut o = npt_chain_baro(st, st->ke_carry, 0.0, false, (double)T, (double)P, d.natoms,
                                   hdt, dt);
        st->vscale = (float)((double)st->v_pending * o.nh_scale * o.baro_scale);
        st->exp_vW = (float)o.exp_vW;
        double h = 0.5 * o.veps_W_dt;
        st->fpos = (float)(dt * exp(h) * sinch(h));
        st->L = (float)o.L;
        st->invL = 1.0f / st->L;
    }
    st->pre_ready = 1;
}

// Push-off speed limit (the displacement cap of a soft push-off): rescales v
// so that |v| <= vlim.
__device__ inline float4 limit_speed(float4 v, float vlim) {
    const float s2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (s2 > vlim * vlim) {
        const float f = vlim * rsqrtf(s2);
        v.x *= f; v.y *= f; v.z *= f;
    }
    return v;
}

// ---------------------------------------------------------------------------
// pre-force: kick, drift, (O), wrap, image, skin displacement
// (src/integrate/langevin.cu:34-94, src/integrate/nose_hoover.cu:606-757,
//  src/neighbor/skin_trigger.cuh:5-16)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_pre(const RepDesc* __restrict__ descs,
                                             const Tile* __restrict__ tiles, int ntl, int* __restrict__ nrt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *nrt = 0;  // k_decide lists this step's rebuilding tiles
    const int w = wave_in_grid();
    if (w >= ntl) return;
    const Tile t = tiles[w];
    const RepDesc& d = descs[t.rep];
    RepState* st = d.st;
    if (!st->active || st->err) return;
    const int i = t.lt * kTile + lane_id();
    float d2 = 0.0f;
    int bad = 0;
    if (i < d.natoms) {
        float4 r = d.pos[i], v = d.vel[i];
        float4 ref = d.pos_ref[i];
        const float m = d.mass[type_of(r.w)];
        const float hdtm = d.half_dt / m;
        const float L = st->L, invL = st->invL;
        if (d.ensemble == 0) {
            if (!st->kicked) {  // else the last force pass gave this half kick
                const float4 f = d.force[i];
                // explicit FMAs: the force kernel's kick ahead computes the same bits
                v.x = fmaf(hdtm, f.x, v.x); v.y = fmaf(hdtm, f.y, v.y); v.z = fmaf(hdtm, f.z, v.z);
                if (d.vlim > 0.0f) v = limit_speed(v, d.vlim);
            }
            r.x += d.half_dt * v.x; r.y += d.half_dt * v.y; r.z += d.half_dt * v.z;
            const float3 z = normal3(d.seed_lo, d.seed_hi, (uint32_t)d.gid[i], st->step + 1);
            const float ns = d.c2 * sqrtf(st->kT / m);
            v.x = d.c1 * v.x + ns * z.x; v.y = d.c1 * v.y + ns * z.y; v.z = d.c1 * v.z + ns * z.z;
            if (d.vlim > 0.0f) v = limit_speed(v, d.vlim);
            r.x += d.half_dt * v.x; r.y += d.half_dt * v.y; r.z += d.half_dt * v.z;
        } else {
            const float sc = st->vscale;
            const float4 f = d.force[i];
            v.x *= sc; v.y *= sc; v.z *= sc;
            v.x += hdtm * f.x; v.y += hdtm * f.y; v.z += hdtm * f.z;
            if (d.ensemble == 1) {
                r.x += d.dt * v.x; r.y += d.dt * v.y; r.z += d.dt * v.z;
            } else {
                const float ew = st->exp_vW, fp = st->fpos;
                r.x = r.x * ew + v.x * fp; r.y = r.y * ew + v.y * fp; r.z = r.z * ew + v.z * fp;
                ref.x *= ew; ref.y *= ew; ref.z *= ew;
                d.pos_ref[i] = ref;
            }
        }
        const int ix = (int)floorf(r.x * invL), iy = (int)floorf(r.y * invL),
                  iz = (int)floorf(r.z * invL);
        r.x -= (float)ix * L; r.y -= (float)iy * L; r.z -= (float)iz * L;

        float dx = r.x - ref.x, dy = r.y - ref.y, dz = r.z - ref.z;
        dx -= L * roundf(dx * invL); dy -= L * roundf(dy * invL); dz -= L * roundf(dz * invL);
        d2 = dx * dx + dy * dy + dz * dz;
        bad = !(isfinite(r.x) && isfinite(r.y) && isfinite(r.z) && isfinite(v.x) &&
                isfinite(v.y) && isfinite(v.z));
        d.pos[i] = r;
        d.vel[i] = v;
        if (ix | iy | iz) {  // image counts change only on a wrap
            int4 im = d.img[i];
            im.x += ix; im.y += iy; im.z += iz;
            d.img[i] = im;
        }
        const double sc = 4294967296.0 / (double)L;
        d.ipos[i] = make_int4((int)fix_coord(r.x, sc), (int)fix_coord(r.y, sc), (int)fix_coord(r.z, sc),
                              __float_as_int(r.w));
    }
    // the tile's largest squared displacement since the last rebuild (-1: a
    // non-finite position or velocity), reduced per replica by k_decide; a
    // per-tile record instead of an atomic on one word per replica, which
    // serializes a large replica's waves
    const float m = wave_max(bad ? 0.0f : d2);
    const int b = wave_or(bad);
    if (lane_id() == 0) d.tdisp[t.lt] = b ? -1.0f : m;
}

// ---------------------------------------------------------------------------
// per-replica rebuild decision (src/main.cu:486-496, skin_trigger.cuh:18-33)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_decide(const RepDesc* __restrict__ descs,
                                                 const int* __restrict__ reps, int nr,
                                                 const int* __restrict__ rep_t0, int* __restrict__ rlist,
                                                 int* __restrict__ nrt) {
    __shared__ float s_max[16];
    __shared__ int s_bad[16];
    __shared__ int s_base;
    const RepDesc& d = descs[reps[blockIdx.x]];
    RepState* st = d.st;
    if (!st->active) return;
    float m = 0.0f;
    int bad = 0;
    for (int tl = threadIdx.x; tl < d.ntiles; tl += blockDim.x) {
        const float x = d.tdisp[tl];
        bad |= (x < 0.0f);
        m = fmaxf(m, x);
    }
    m = wave_max(m);
    bad = wave_or(bad);
    if (lane_id() == 0) {
        s_max[threadIdx.x >> 6] = m;
        s_bad[threadIdx.x >> 6] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_base = -1;
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            m = fmaxf(m, s_max[w]);
            bad |= s_bad[w];
        }
        m = fmaxf(m, s_max[0]);
        bad |= s_bad[0];
        if (bad) st->err |= ERR_NAN;
        if (st->err) {
            st->active = 0;
            st->rebuild = 0;
        } else {
            // k_pre of this step placed the positions in the box st->L
            st->Lp = st->L;
            st->invLp = st->invL;
            set_grid_consts(st, d.rc2);
            const float md = fmaxf(__int_as_float(st->maxdisp_bits), m);  // largest since the last rebuild
            const int since = st->since_rebuild + 1;
            const int rb = (md > d.half_skin2) || (since >= kRebuildEvery);
            st->rebuild = rb;
            if (rb) {
                st->since_rebuild = 0;
                st->maxdisp_bits = 0;
                st->nrebuilds++;
                set_grid(d, st);
                s_base = atomicAdd(nrt, d.ntiles);  // this replica's run of the rebuild tile list
            } else {
                st->since_rebuild = since;
                st->maxdisp_bits = __float_as_int(md);
            }
        }
    }
    __syncthreads();
    // the replica's tiles in the rebuild tile list (whole groups, in any
    // order of replicas: every consumer treats each group alone)
    const int base = s_base;
    if (base >= 0)
        for (int k = threadIdx.x; k < d.ntiles; k += blockDim.x) rlist[base + k] = rep_t0[blockIdx.x] + k;
}

// Radius of gyration of each chain (src/analysis/rg.cu:3-69): unwrapped
// positions x + image * L over the chain's beads (global-id order, through
// slot_of), centre of mass, then sqrt of the mean squared distance from it.
// One wave per chain, FP64 sums by wave shuffles.
__global__ void __launch_bounds__(64) k_rg(const RepDesc* __restrict__ d, double* __restrict__ out) {
    const int c = blockIdx.x;
    if (c >= d->nchains) return;
    const int g0 = d->t_choff[c], g1 = d->t_choff[c + 1], len = g1 - g0;
    const double L = (double)d->st->Lp;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int g = g0 + lane_id(); g < g1; g += 64) {
        const int q = d->slot_of[g];
        const float4 p = d->pos[q];
        const int4 im = d->img[q];
        sx += (double)p.x + im.x * L; sy += (double)p.y + im.y * L; sz += (double)p.z + im.z * L;
    }
    const double cx = wave_sum(sx) / len, cy = wave_sum(sy) / len, cz = wave_sum(sz) / len;
    double r2 = 0.0;
    for (int g = g0 + lane_id(); g < g1; g += 64) {
        const int q = d->slot_of[g];
        const float4 p = d->pos[q];
        const int4 im = d->img[q];
        const double dx = (double)p.x + im.x * L - cx, dy = (double)p.y + im.y * L - cy,
                     dz = (double)p.z + im.z * L - cz;
        r2 += dx * dx + dy * dy + dz * dz;
    }
    r2 = wave_sum(r2);
    if (lane_id() == 0) out[c] = sqrt(r2 / len);
}

// Mean of the per-chain values k_rg wrote to vals, summed in chain order by
// one lane (the order the host sum of the per-chain read uses).
__global__ void __launch_bounds__(64) k_rg_mean(const RepDesc* __restrict__ d, const double* __restrict__ vals,
                                                double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    const int nc = d->nchains;
    double s = 0.0;
    for (int c = 0; c < nc; c++) s += vals[c];
    out[0] = nc > 0 ? s / nc : 0.0;
}

// Status of each listed replica for one host read per batch.
__global__ void k_status(const RepDesc* __restrict__ descs, const int* __restrict__ reps, int nr,
                         int4* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr) return;
    const RepState* st = descs[reps[k]].st;
    const uint64_t s = (uint64_t)st->step;
    out[k] = make_int4(st->err, st->active, (int)(uint32_t)s, (int)(uint32_t)(s >> 32));
}

// Steps since each listed replica's last list build (-1: inactive), for the
// host's choice of step structure.
__global__ void k_phase(const RepDesc* __restrict__ descs, const int* __restrict__ reps, int nr,
                        int* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr) return;
    const RepState* st = descs[reps[k]].st;
    out[k] = st->active ? st->since_rebuild : -1;
}

// Host-requested state changes.
__global__ void k_control(const RepDesc* __restrict__ descs, const int* __restrict__ reps, int nr,
                          int op, int64_t arg) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr) return;
    const RepDesc& d = descs[reps[k]];
    RepState* st = d.st;
    if (op == CTL_RUN) {
        st->stop = st->step + arg;
        st->active = (st->err == 0 && arg > 0) ? 1 : 0;
    } else if (op == CTL_FORCE_REBUILD) {
        st->rebuild = 1;
        st->since_rebuild = 0;
        st->maxdisp_bits = 0;
        st->nrebuilds++;
        set_grid(d, st);
    }
}

// ---------------------------------------------------------------------------
// Morton-ordered cell sort.  Cells have side >= rc+skin on the reference grid;
// the cell key is the 3-D Morton interleave of the cell coordinates, so the
// slot order is a Morton curve over cells, and within a cell atoms are ordered
// by global id (deterministic).
// ---------------------------------------------------------------------------
// The rebuild kernels walk the batch's compact list of rebuilding tiles
// (k_rebuild_list), so their cost follows the replicas that rebuild.
#define REBUILD_PROLOGUE                                      \
    const Tile t = tiles[rlist[v]];                           \
    const RepDesc& d = descs[t.rep];                          \
    RepState* st = d.st;                                      \
    (void)st;                                                 \
    const int i = t.lt * kTile + lane_id();
#define REBUILD_LOOP for (int v = wave_in_grid(), nv = *nrt; v < nv; v += waves_in_grid())
#define REBUILD_ARGS const RepDesc* __restrict__ descs, const Tile* __restrict__ tiles, \
                     const int* __restrict__ rlist, const int* __restrict__ nrt

// Compact list of the tiles of rebuilding replicas (all replicas if `always`):
// one workgroup scans the per-replica tile counts and writes tile indices.
__global__ void __launch_bounds__(1024) k_rebuild_list(const RepDesc* __restrict__ descs,
                                                       const int* __restrict__ reps,
                                                       const int* __restrict__ rep_t0, int nr,
                                                       int always, int* __restrict__ rlist,
                                                       int* __restrict__ nrt) {
    __shared__ int wsum[16];
    __shared__ int s_off[kMaxReplicas];  // exclusive prefix of the listed tile counts
    __shared__ int s_carry;
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    // one replica per thread and pass, a block prefix per pass of 1024
    for (int r0 = 0; r0 < nr; r0 += 1024) {
        const int r = r0 + threadIdx.x;
        int c = 0;
        if (r < nr) {
            const RepDesc& d = descs[reps[r]];
            const RepState* st = d.st;
            if (always || (st->active && st->rebuild)) c = d.ntiles;
        }
        int x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        int before = 0, all = 0;
        for (int w = 0; w < 16; w++) {
            const int ws = wsum[w];
            before += (w < wv) ? ws : 0;
            all += ws;
        }
        const int carry = s_carry;
        if (r < nr) s_off[r] = carry + before + x - c;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + all;
        __syncthreads();
    }
    const int total = s_carry;
    if (threadIdx.x == 0) *nrt = total;
    // each wave writes the tiles of replicas wv, wv + 16, ... (coalesced runs)
    for (int r = wv; r < nr; r += 16) {
        const int o = s_off[r];
        const int n = ((r + 1 < nr) ? s_off[r + 1] : total) - o;
        if (n == 0) continue;
        const int t0 = rep_t0[r];
        for (int v = lane; v < n; v += 64) rlist[o + v] = t0 + v;
    }
}

// Cell key of each atom and a unique position inside its cell.  The lanes of
// a wave (Morton-sorted atoms) fall in a few cells: one atomic per distinct
// key and wave reserves the positions, handed out in lane order.
__global__ void __launch_bounds__(256) k_bin(REBUILD_ARGS) {
    REBUILD_LOOP {
        REBUILD_PROLOGUE
        const bool valid = i < d.natoms;
        uint32_t key = 0xffffffffu;
        if (valid) {
            const float4 r = d.pos[i];
            const int nx = st->nx;
            const float ics = st->inv_cs;
            const int cx = cell_coord(r.x, ics, nx);
            key = cell_key(cx, cell_coord(r.y, ics, nx), cell_coord(r.z, ics, nx));
            d.cell_of[i] = (int)key;
            // the within-cell sort key (k_rank): the x grid coordinate
            // floor(x 2^32 / L) modulo 2^32, relative to the cell's lower bound
            // floor(cx 2^32 / nx)
            const unsigned lo = (unsigned)(((unsigned long long)cx << 32) / (unsigned)nx);
            d.sortx[i] = (int)(fix_coord(r.x, 4294967296.0 / (double)st->Lp) - lo);
        }
        unsigned long long todo = __ballot(valid);
        while (todo) {  // wave-uniform: one pass per distinct key
            const int lead = __builtin_ctzll(todo);
            const uint32_t k = __shfl(key, lead, 64);
            const unsigned long long same = __ballot(valid && key == k);
            int base = 0;
            if (lane_id() == lead) base = atomicAdd(&d.cell_cnt[k], __popcll(same));
            base = __shfl(base, lead, 64);
            if (valid && key == k) d.rank[i] = base + lane_prefix(same);
            todo &= ~same;
        }
    }
}

// Exclusive scan of the cell counts, one 1024-thread workgroup per replica;
// the counts are cleared for the next rebuild.
__global__ void __launch_bounds__(1024) k_scan(const RepDesc* __restrict__ descs,
                                               const int* __restrict__ reps, int nr, int always) {
    // chunks of 1024 x 4 keys: thread t holds keys 4t..4t+3 of the chunk
    // (one 16-byte load; a wave covers 1 KB contiguous), block-wide scan,
    // then the running offset carries to the next chunk
    __shared__ int wsum[16];
    __shared__ int carry;
    const RepDesc& d = descs[reps[blockIdx.x]];
    RepState* st = d.st;
    if (!always && !(st->active && st->rebuild)) return;
    const int nkeys = 1 << (3 * st->cbits);
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nkeys; base += 4 * 1024) {
        const int c0 = base + 4 * threadIdx.x;
        int4 v = make_int4(0, 0, 0, 0);
        if (c0 + 3 < nkeys) {
            v = *(const int4*)(d.cell_cnt + c0);
        } else {
            if (c0 < nkeys) v.x = d.cell_cnt[c0];
            if (c0 + 1 < nkeys) v.y = d.cell_cnt[c0 + 1];
            if (c0 + 2 < nkeys) v.z = d.cell_cnt[c0 + 2];
        }
        const int local = v.x + v.y + v.z + v.w;
        int x = local;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = carry;
            for (int k = 0; k < 16; k++) { const int t = wsum[k]; wsum[k] = acc; acc += t; }
            carry = acc;
        }
        __syncthreads();
        const int run = wsum[wv] + x - local;  // exclusive prefix of this thread's keys
        if (c0 + 3 < nkeys) {
            *(int4*)(d.cell_start + c0) = make_int4(run, run + v.x, run + v.x + v.y, run + v.x + v.y + v.z);
            *(int4*)(d.cell_cnt + c0) = make_int4(0, 0, 0, 0);
        } else {
            if (c0 < nkeys) { d.cell_start[c0] = run; d.cell_cnt[c0] = 0; }
            if (c0 + 1 < nkeys) { d.cell_start[c0 + 1] = run + v.x; d.cell_cnt[c0 + 1] = 0; }
            if (c0 + 2 < nkeys) { d.cell_start[c0 + 2] = run + v.x + v.y; d.cell_cnt[c0 + 2] = 0; }
        }
        __syncthreads();  // wsum is rewritten by the next chunk
    }
    if (threadIdx.x == 0) d.cell_start[nkeys] = carry;
}

__global__ void __launch_bounds__(256) k_scatter(REBUILD_ARGS) {
    REBUILD_LOOP {
        REBUILD_PROLOGUE
        if (i < d.natoms) {
            const int dst = d.cell_start[d.cell_of[i]] + d.rank[i];
            d.order[dst] = i;
            d.tmp_pos[i] = d.pos[i];
            d.tmp_vel[i] = d.vel[i];
            d.tmp_img[i] = d.img[i];
        }
    }
}

// Rank inside the cell by x, then global id: a deterministic order
// independent of the atomic arrival order of k_bin, in which the list build
// finds the atoms of a row of cells within reach along x by binary search.
// x is the grid coordinate floor(x 2^32 / L) taken modulo 2^32 relative to the
// cell's lower bound floor(cx 2^32 / nx) (k_bin's sortx): the list build
// compares grid coordinates modulo 2^32, and a position a rounding step
// outside [0, L) (at -0 or L itself) stands at the cell's end where the grid
// coordinate puts it.
__global__ void __launch_bounds__(256) k_rank(REBUILD_ARGS) {
    REBUILD_LOOP {
        REBUILD_PROLOGUE
        if (i < d.natoms) {
            const int src = d.order[i];
            const int key = d.cell_of[src];
            const int s = d.cell_start[key], e = d.cell_start[key + 1];
            const int g = d.tmp_img[src].w;
            const int rx = d.sortx[src];
            int r = 0;
            for (int q = s; q < e; q++) {
                const int o = d.order[q];
                const int ro = d.sortx[o];
                r += (ro < rx) || (ro == rx && d.tmp_img[o].w < g);
            }
            d.order2[s + r] = src;
        }
    }
}

__global__ void __launch_bounds__(256) k_gather(REBUILD_ARGS) {
    REBUILD_LOOP {
        REBUILD_PROLOGUE
        if (i < d.natoms) {
            const int src = d.order2[i];
            const float4 p = d.tmp_pos[src];
            const int4 im = d.tmp_img[src];
            d.pos[i] = p;
            d.pos_ref[i] = p;
            d.vel[i] = d.tmp_vel[src];
            d.img[i] = im;
            d.gid[i] = im.w;
            d.slot_of[im.w] = i;
            const double sc = 4294967296.0 / (double)st->Lp;
            d.ipos[i] = make_int4((int)fix_coord(p.x, sc), (int)fix_coord(p.y, sc), (int)fix_coord(p.z, sc),
                                  __float_as_int(p.w));
        }
    }
}

// ---------------------------------------------------------------------------
// Verlet list + bonded slot tables (src/neighbor/verlet_list.cu:108-335).
// Full list, r^2 < (rc+skin)^2 on the reference FP32 distance, i != j, 1-2
// bonded partners excluded; stored tile-major nbr[tile][k][lane] so the force
// kernel streams one contiguous row per neighbor index.  A count above kcap is
// reported as ERR_NBR_OVERFLOW, never silently truncated.
//
// One workgroup of kGroupTiles waves per group of tiles.  The group's atoms lie
// in a few Morton-consecutive cells; the union of those cells' 27-neighborhoods
// (typically 90-140 cells, about 2000-2900 atoms) is staged in LDS with
// coalesced loads (the top 16 bits of each grid coordinate: a prefilter that
// sends the few candidates near the radius to the exact FP32 test), and each
// lane scans its own 27 cells from LDS in
// (dz, dy, dx) order, candidates in ascending slot order, exactly as the
// global-memory scan that serves groups whose union does not fit.  Both give
// the same rows.  The atoms the group's rows reference (its neighbor set, about
// 1400) are then numbered in slot order and listed in gstage; the rows are
// rewritten to those atoms' byte offsets in the force kernel's LDS copy of the
// set (or to slots when the set exceeds kGroupStage).
// ---------------------------------------------------------------------------
constexpr int kMaxWin = 1024;      // cells of a group's neighbor window (bitmap words <= 32)
constexpr int kPlaceBand = 16;     // row entries moved per pass of the placement

// cell coordinate offset by o in -1..1 with periodic wrap
__device__ inline int wrap_cell(int c, int o, int nx) {
    c += o;
    c += (c < 0) ? nx : 0;
    c -= (c >= nx) ? nx : 0;
    return c;
}

// Appends the candidates of mask m (bit k: slot q0 + k) in slot order,
// skipping the atom itself and its bonded partners.  Entry k is stored as
// ebase + k (a staged index or a slot) and, with `used`, marks used[ebase + k].
// W16 rows hold 16-bit entries (replicas of at most 65536 slots), others 32-bit.
// The bits of candidates q0 + k that are the atom itself or a bonded partner.
__device__ __forceinline__ unsigned excluded_mask(unsigned m, int q0, int i, const int (&ex)[4], int nb, int bcap,
                                                  const FMD_GLOBAL int* bt) {
    unsigned x = 0u;
    auto drop = [&](int e) {
        const unsigned o = (unsigned)(e - q0);
        x |= (o < (unsigned)kBuildChunk) ? (1u << o) : 0u;
    };
    drop(i); drop(ex[0]); drop(ex[1]); drop(ex[2]); drop(ex[3]);
    for (int b = 4; b < nb && b < bcap; b++) drop(bt[b * kTile] & kSlotMask);
    return m & x;
}

template <bool W16>
__device__ __forceinline__ void append_chunk(unsigned m, int q0, int i, const int (&ex)[4], int nb, int bcap,
                                             const FMD_GLOBAL int* bt, FMD_GLOBAL int* out,
                                             FMD_GLOBAL unsigned short* out16, int kcap, int& cnt, int ebase,
                                             unsigned char* used) {
    m &= ~excluded_mask(m, q0, i, ex, nb, bcap, bt);
    // one pass per set bit, lowest first: a wave runs as many passes as its
    // fullest lane has pairs in the chunk (about 4 of 8)
    while (m) {
        const int k = __builtin_ctz(m);
        m &= m - 1u;
        if (cnt < kcap) {
            if (W16) out16[cnt * kTile] = (unsigned short)(ebase + k);
            else out[cnt * kTile] = ebase + k;
            if (used) used[ebase + k] = 1;
        }
        cnt++;
    }
}

// One chunk of kBuildChunk candidates (slots q0 + k, k < n) by the reference
// FP32 distance: the tests give a bit mask, then the set bits are appended.
template <bool W16>
__device__ __forceinline__ void build_chunk(const float4 (&pj)[kBuildChunk], int q0, int n, int i, float4 pi,
                                            float L, float invL, float rcs2, const int (&ex)[4], int nb,
                                            int bcap, const FMD_GLOBAL int* bt, FMD_GLOBAL int* out,
                                            FMD_GLOBAL unsigned short* out16, int kcap, int& cnt, int ebase,
                                            unsigned char* used) {
    unsigned m = 0;
#pragma unroll
    for (int k = 0; k < kBuildChunk; k++) {
        const float r2 = dist2_ref(pi, pj[k], L, invL);  // every k: no per-candidate branch
        m |= ((k < n) & (r2 < rcs2)) ? (1u << k) : 0u;
    }
    append_chunk<W16>(m, q0, i, ex, nb, bcap, bt, out, out16, kcap, cnt, ebase, used);
}

// The same decision with a coarse fixed-point prefilter.  The staged
// positions keep the top 16 bits of each grid coordinate, x and y packed in
// one word and z in the low half of another, so one packed 16-bit subtract per
// word gives the minimum-image differences (modulo 2^16, as the full grid
// wraps modulo 2^32) and two 16-bit dot products give their squared length
// qc.  Truncation moves each coordinate by less than one coarse unit, so
// |sqrt(qc) - r| < sqrt(3) units: qc below clo is certainly inside the list
// radius and qc above chi certainly outside (both bounds also allow for the
// band the exact FP32 test may differ by).  The few candidates in between --
// about 1e-3 sigma either side of the radius -- are decided by the reference
// FP32 formula on their positions.
template <bool B> struct BoolC { static constexpr bool value = B; };
typedef short fmd_s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short fmd_u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ fmd_s16x2 coarse_diff(unsigned a, unsigned b) {  // two 16-bit lanes, modulo 2^16
    return __builtin_bit_cast(fmd_s16x2, __builtin_bit_cast(fmd_u16x2, a) - __builtin_bit_cast(fmd_u16x2, b));
}
// qc + acc for the two packed coordinate words of each atom, saturating (a
// sum past INT_MAX is far outside any radius); acc enters the first product
__device__ __forceinline__ int coarse_t(unsigned axy, unsigned az, unsigned bxy, unsigned bz, int acc) {
    const fmd_s16x2 dxy = coarse_diff(axy, bxy), dz = coarse_diff(az, bz);
    return __builtin_amdgcn_sdot2(dz, dz, __builtin_amdgcn_sdot2(dxy, dxy, acc, true), true);
}
__device__ __forceinline__ unsigned coarse_xy(int4 p) {
    return ((unsigned)p.x >> 16) | ((unsigned)p.y & 0xFFFF0000u);
}
__device__ __forceinline__ unsigned coarse_z(int4 p) { return (unsigned)p.z >> 16; }
// The lowest r set bits of m (none for r <= 0).
__device__ inline unsigned lowest_bits(unsigned m, int r) {
    unsigned keep = 0u;
    for (int j = 0; j < r && m; j++) {
        const unsigned b = m & (0u - m);
        keep |= b;
        m ^= b;
    }
    return keep;
}
// A candidate the prefilter leaves undecided needs the reference distance on
// its float position, a global load.  The first such candidate of a lane is
// appended as inside and recorded in pend (row position << 12 | staged index);
// after the scan the lane decides it, with one load round trip for the whole
// wave, and turns it into the atom's own entry (which every consumer skips) if
// it lies outside.  Further undecided candidates of the lane are decided at
// once.  Candidates are staged indices q0 + k (k < n, n >= 1) with coarse
// positions jp[k] = (x | y << 16, z); x holds the bits of the candidates to
// leave out (the atom itself and its bonded partners; bits from kBuildChunk
// up are ignored), bw = chi - clo is the width of the undecided band, slot_of
// maps a staged index to its slot, and used[q] marks atom q as in a row.
template <bool W16, class SlotOf>
__device__ __forceinline__ void build_chunk_q(const uint2 (&jp)[kBuildChunk], unsigned pxy, unsigned pz, int clo,
                                              unsigned bw, const FMD_GLOBAL float4* gpos, int q0, int n, float4 pi,
                                              float L, float invL, float rcs2, unsigned x, FMD_GLOBAL int* out,
                                              FMD_GLOBAL unsigned short* out16, int kcap, int& cnt, unsigned char* used,
                                              int& pend, SlotOf slot_of) {
    // t = q - clo is negative for a candidate certainly inside.  Its sign bit
    // is shifted into m from the last candidate down (alignbit of (m, t) by 31
    // is m << 1 | t >> 31), so bit k of m is candidate k; the smallest t as an
    // unsigned number is the nearest candidate not certainly inside, and only
    // when it lies in the band [0, bw] are the band's candidates found one by
    // one.
    unsigned m = 0u, umin = ~0u;
    int t[kBuildChunk];
#pragma unroll
    for (int k = kBuildChunk - 1; k >= 0; k--) {
        t[k] = coarse_t(pxy, pz, jp[k].x, jp[k].y, -clo);
        m = __builtin_amdgcn_alignbit(m, (unsigned)t[k], 31);
        umin = min(umin, (unsigned)t[k]);
    }
    const unsigned valid = (1u << min(n, kBuildChunk)) - 1u;  // k < n
    m &= valid & ~x;
    if (umin <= bw) {
        unsigned band = 0u;
#pragma unroll
        for (int k = 0; k < kBuildChunk; k++) band |= ((unsigned)t[k] <= bw) ? (1u << k) : 0u;
        band &= valid & ~x;
        if (band && pend < 0) {  // defer the first: inside for now, its row position recorded
            const int k = __builtin_ctz(band);
            band &= band - 1u;
            pend = ((cnt + __popc(m & ((1u << k) - 1u))) << 12) | (q0 + k);
            m |= 1u << k;
        }
        for (int k = 0; k < kBuildChunk; k++)
            if ((band >> k) & 1u)
                m |= (dist2_ref(pi, gpos[slot_of(q0 + k)], L, invL) < rcs2) ? (1u << k) : 0u;
    }
    // the set bits in order, one pass per bit (a wave runs as many passes as
    // its fullest lane has pairs in the chunk); entries past the row capacity
    // are counted, not stored
    const int p = __popc(m);
    if (cnt + p > kcap) m = lowest_bits(m, kcap - cnt);
    unsigned off = (unsigned)cnt * kTile;
    cnt += p;
    while (m) {
        const int k = __builtin_ctz(m);
        m &= m - 1u;
        if (W16) out16[off] = (unsigned short)(q0 + k);
        else out[off] = q0 + k;
        used[q0 + k] = 1;
        off += kTile;
    }
}

// Block-wide exclusive prefix of one int per thread (kGroupSlots threads).
__device__ inline int block_excl_scan(int x, int* s_wsum, int* total) {
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    int v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) s_wsum[wv] = v;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kGroupTiles; w++) {
        const int ws = s_wsum[w];
        before += (w < wv) ? ws : 0;
        all += ws;
    }
    __syncthreads();
    *total = all;
    return before + v - x;
}

// At most 128 VGPRs (four waves per SIMD, four workgroups per CU with the
// LDS below): the rebuilding replicas' groups fit in fewer rounds.
__global__ void __launch_bounds__(kGroupSlots) __attribute__((amdgpu_waves_per_eu(4, 8))) k_build(REBUILD_ARGS) {
    // staged union, coarse fixed point: x | y << 16 and z of each atom side by
    // side (one chunk of padding: a cell's last chunk reads whole)
    __shared__ uint2 s_pj[kBuildStage + kBuildChunk];
    __shared__ unsigned short s_seg[kBuildStage];  // staged atom -> union cell
    __shared__ unsigned char s_used[kBuildStage];  // staged atom is in a row of the group
    __shared__ int4 s_cell[kMaxUnion];             // (first slot, count, LDS offset, key)
    __shared__ unsigned int s_bits[kMaxWin / 32];   // union cells of the window
    __shared__ unsigned char s_wmap[kMaxWin];       // window cell -> union index
    __shared__ int s_red[6][kGroupTiles];           // per-wave window bounds
    __shared__ int s_wsum[kGroupTiles];
    __shared__ int s_c0[3], s_nu, s_nstage;
    __shared__ int s_len[kGroupSlots];              // padded row length of each atom of the group
    unsigned short* s_cidx = (unsigned short*)s_pj;  // after the scan: compact index of a used atom
    // the wave index as a scalar: each wave's tile and replica fields are uniform
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = lane_id(), tid = threadIdx.x;
    for (int gb = blockIdx.x, ng = *nrt / kGroupTiles; gb < ng; gb += gridDim.x) {
        const Tile t = tiles[rlist[gb * kGroupTiles + wv]];
        const RepDesc& d = descs[t.rep];
        RepState* st = d.st;
        const int i = t.lt * kTile + lane;
        const int grp = t.lt / kGroupTiles;
        const bool valid = i < d.natoms;
        int err = 0;
        // bonded partners (slots), the exclusion set and the bonded tables
        int ex[4] = {-1, -1, -1, -1};
        int nb = 0;
        float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid) {
            pi = gview(d.pos)[i];
            const int g = gview(d.img)[i].w;
            const int b0 = gview(d.t_boff)[g], b1 = gview(d.t_boff)[g + 1];
            for (int e = b0; e < b1; e++) {
                const int ent = gview(d.t_blst)[e];
                const int ps = gview(d.slot_of)[ent & kSlotMask];
                if (nb < d.bcap) gmut(d.bnd)[((size_t)t.lt * d.bcap + nb) * kTile + lane] = ps | (ent & ~kSlotMask);
                if (nb < 4) ex[nb] = ps;
                nb++;
            }
            if (nb > d.bcap) err |= ERR_BONDED;
            int na = 0;
            const int a0 = gview(d.t_aoff)[g], a1 = gview(d.t_aoff)[g + 1];
            for (int e = a0; e < a1; e++) {
                const int4 A = gview(d.t_alst)[e];
                const int role = (A.x == g) ? 0 : ((A.y == g) ? 1 : 2);
                const int o1 = (role == 0) ? A.y : A.x;
                const int o2 = (role == 2) ? A.y : A.z;
                const int s1 = gview(d.slot_of)[o1], s2 = gview(d.slot_of)[o2];
                if (na < d.acap)
                    gmut(d.ang)[((size_t)t.lt * d.acap + na) * kTile + lane] =
                        make_int2(s1 | (role << kSlotBits), s2 | (A.w << kSlotBits));
                na++;
            }
            if (na > d.acap) err |= ERR_BONDED;
            gmut(d.bhdr)[i] = min(nb, d.bcap) | (min(na, d.acap) << 8);
        }
        const FMD_GLOBAL int* bt = gview(d.bnd) + (size_t)t.lt * d.bcap * kTile + lane;
        FMD_GLOBAL int* out = gmut(d.nbr) + (size_t)t.lt * d.kcap * kTile + lane;
        // 16-bit rows are built in the upper half of the tile's row area (which
        // holds 2 kcap 16-bit entries per lane) and moved to the lower half by
        // the placement below; 32-bit rows are built in place
        FMD_GLOBAL unsigned short* out16 =
            (FMD_GLOBAL unsigned short*)(gmut(d.nbr) + (size_t)t.lt * d.kcap * kTile) + (size_t)d.kcap * kTile + lane;
        const int kcap = d.kcap;
        const float rcs2 = d.rcs2;
        const int nx = st->nx;
        const float ics = st->inv_cs;
        const float L = st->Lp, invL = st->invLp;
        const FMD_GLOBAL float4* gpos = gview(d.pos);
        const FMD_GLOBAL int* cstart = gview(d.cell_start);
        int cx = 0, cy = 0, cz = 0;
        if (valid) {
            cx = cell_coord(pi.x, ics, nx); cy = cell_coord(pi.y, ics, nx); cz = cell_coord(pi.z, ics, nx);
        }
        // The group's cells lie in a small box of cell coordinates (taken
        // relative to the cell of the group's first atom, which is always a real
        // atom, modulo the box).  Their 27-neighborhoods fall in that box grown
        // by one cell per side (the whole axis if that wraps): a window of at
        // most kMaxWin cells.  A bitmap over the window marks the union, whose
        // cells are numbered in window order.
        if (tid == 0) { s_c0[0] = cx; s_c0[1] = cy; s_c0[2] = cz; }
        __syncthreads();
        const int h = nx >> 1;
        auto rel = [&](int c, int c0) {
            int r = c - c0;
            r += (r < -h) ? nx : 0;
            r -= (r >= nx - h) ? nx : 0;
            return r;
        };
        const int c0x = s_c0[0], c0y = s_c0[1], c0z = s_c0[2];
        const int rx = valid ? rel(cx, c0x) : 0, ry = valid ? rel(cy, c0y) : 0, rz = valid ? rel(cz, c0z) : 0;
        {
            const int lo_x = wave_min_i(valid ? rx : INT_MAX), hi_x = wave_max_i(valid ? rx : INT_MIN);
            const int lo_y = wave_min_i(valid ? ry : INT_MAX), hi_y = wave_max_i(valid ? ry : INT_MIN);
            const int lo_z = wave_min_i(valid ? rz : INT_MAX), hi_z = wave_max_i(valid ? rz : INT_MIN);
            if (lane == 0) {
                s_red[0][wv] = lo_x; s_red[1][wv] = hi_x; s_red[2][wv] = lo_y;
                s_red[3][wv] = hi_y; s_red[4][wv] = lo_z; s_red[5][wv] = hi_z;
            }
        }
        __syncthreads();
        int sx, sy, sz, ex_, ey_, ez_;
        auto window = [&](int a, int& s0, int& e0) {
            int lo = INT_MAX, hi = INT_MIN;
            for (int w = 0; w < kGroupTiles; w++) { lo = min(lo, s_red[2 * a][w]); hi = max(hi, s_red[2 * a + 1][w]); }
            s0 = lo - 1;
            e0 = min(hi - lo + 3, nx);  // the whole axis once the grown box wraps
        };
        window(0, sx, ex_);
        window(1, sy, ey_);
        window(2, sz, ez_);
        // window cell of this atom's neighbor cell c in (dz, dy, dx) order; the
        // offset from the window start is >= 0 and, on an axis the window
        // wraps (extent nx), taken modulo nx
        auto wmod = [](int v, int e) {  // v in [0, e + 1]: v mod e (e >= 1)
            v -= (v >= e) ? e : 0;
            v -= (v >= e) ? e : 0;
            return v;
        };
        auto win_of = [&](int c) {
            const int wx = wmod(rx + c % 3 - 1 - sx, ex_), wy = wmod(ry + (c / 3) % 3 - 1 - sy, ey_),
                      wz = wmod(rz + c / 9 - 1 - sz, ez_);
            return (wz * ey_ + wy) * ex_ + wx;
        };
        const int W = ex_ * ey_ * ez_;
        // The staged index of a bonded partner (slot ps): its cell is one of the
        // 27 around this atom's (a bond is shorter than a cell), found through
        // the window; -1 if not (the group then keeps slots everywhere).
        auto partner_q = [&](int ps) __attribute__((always_inline)) -> int {
            const float4 pp = gpos[ps];
            int dx = cell_coord(pp.x, ics, nx) - cx, dy = cell_coord(pp.y, ics, nx) - cy,
                dz = cell_coord(pp.z, ics, nx) - cz;
            dx += (dx < -1) ? nx : 0; dx -= (dx > 1) ? nx : 0;
            dy += (dy < -1) ? nx : 0; dy -= (dy > 1) ? nx : 0;
            dz += (dz < -1) ? nx : 0; dz -= (dz > 1) ? nx : 0;
            if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return -1;
            const int u = s_wmap[win_of((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1))];
            if (u >= s_nu) return -1;
            const int4 ce = s_cell[u];
            const int o = ps - ce.x;
            return (o >= 0 && o < ce.y) ? ce.z + o : -1;
        };
        // staged scan: prefilter in fixed point; boxes of fewer than 3 cells per
        // axis (where a pair can sit near half the box) use the global scan
        bool staged = W <= kMaxWin && nx >= 3;  // uniform over the group
        const int4 pf = valid ? (int4)gview(d.ipos)[i] : make_int4(0, 0, 0, 0);
        const unsigned cxy = coarse_xy(pf), cz16 = coarse_z(pf);
        int clo, chi, rb2;
        {
            // radius in coarse units (L / 2^16), widened by the band the exact
            // FP32 test may differ by and by sqrt(3) units of truncation
            const double rc = (double)d.rcs * 65536.0 / (double)L;
            const double band = 3e-5 + 1e-6 * (double)L / (double)d.rcs;  // >> the two forms' difference
            const double rb = rc * sqrt(1.0 + band);  // above the true distance of any pair the FP32 test accepts
            const double rin = rc * sqrt(1.0 - band) - 1.7321 - 0.01, rout = rb + 1.7321 + 0.01;
            clo = rin > 0.0 ? (int)floor(rin * rin) : 0;
            chi = (int)ceil(rout * rout);
            rb2 = (int)fmin(ceil(rb * rb) + 1.0, 2147483647.0);
        }
        int nu = 0, nstage = 0;
        if (staged) {
            for (int j = tid; j < kMaxWin / 32; j += kGroupSlots) s_bits[j] = 0u;
            __syncthreads();
            // each cell of the group marks its 27 neighbors (its first lane in a wave does)
            const int prev = __shfl_up(cx + nx * (cy + nx * cz), 1, 64);
            if (valid && (lane == 0 || prev != cx + nx * (cy + nx * cz)))
                #pragma unroll 1
                for (int c = 0; c < 27; c++) {
                    const int w = win_of(c);
                    atomicOr(&s_bits[w >> 5], 1u << (w & 31));
                }
            __syncthreads();
            if (wv == 0) {
                // number the marked cells in window order
                const int nwords = (W + 31) >> 5;  // <= kMaxWin / 32 <= 64 lanes
                const unsigned bits = (lane < nwords) ? s_bits[lane] : 0u;
                int x = __popc(bits);
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int y = __shfl_up(x, o, 64);
                    if (lane >= o) x += y;
                }
                const int nuw = __shfl(x, 63, 64);
                if (lane == 0) s_nu = nuw;
                if (nuw <= kMaxUnion) {
                    // lane l numbers bit (l & 31) of word j + (l >> 5): two words per pass
                    const int base = x - __popc(bits);  // marked cells in words below this lane's word
                    auto wrapn = [&](int p) {           // p in [-nx, 2 nx)
                        p += (p < 0) ? nx : 0;
                        p -= (p >= nx) ? nx : 0;
                        return p;
                    };
                    for (int j = 0; j < nwords; j += 2) {
                        const int word = j + (lane >> 5), bit = lane & 31;
                        const unsigned wb = __shfl(bits, word & 63, 64);
                        const int wbase = __shfl(base, word & 63, 64);
                        if (word < nwords && ((wb >> bit) & 1u)) {
                            const int w = word * 32 + bit;
                            const int u = wbase + __popc(wb & ((1u << bit) - 1u));
                            const int wx = w % ex_, wy = (w / ex_) % ey_, wz = w / (ex_ * ey_);
                            s_wmap[w] = (unsigned char)u;
                            s_cell[u] = make_int4(0, 0, 0, (int)cell_key(wrapn(c0x + sx + wx),
                                                                         wrapn(c0y + sy + wy),
                                                                         wrapn(c0z + sz + wz)));
                        }
                    }
                    // counts and LDS offsets of the union cells
                    int ns = 0;
                    for (int u0 = 0; u0 < nuw; u0 += 64) {
                        const int u = u0 + lane;
                        int s0 = 0, c = 0;
                        if (u < nuw) {
                            const int k = s_cell[u].w;
                            s0 = cstart[k];
                            c = cstart[k + 1] - s0;
                        }
                        int y = c;
#pragma unroll
                        for (int o = 1; o < 64; o <<= 1) {
                            const int z = __shfl_up(y, o, 64);
                            if (lane >= o) y += z;
                        }
                        if (u < nuw) s_cell[u] = make_int4(s0, c, ns + y - c, s_cell[u].w);
                        ns += __shfl(y, 63, 64);
                    }
                    if (lane == 0) s_nstage = ns;
                }
            }
            __syncthreads();
            nu = s_nu;
            staged = nu <= kMaxUnion;
            if (staged) {
                nstage = s_nstage;
                staged = nstage <= kBuildStage;
            }
        }
        if (staged) {
            for (int u = tid; u < nu; u += kGroupSlots) {
                const int4 ce = s_cell[u];
                for (int j = 0; j < ce.y; j++) s_seg[ce.z + j] = (unsigned short)u;
            }
            for (int q = tid; q < nstage; q += kGroupSlots) s_used[q] = 0;
            __syncthreads();
            const FMD_GLOBAL int4* gip = gview(d.ipos);
            for (int q0 = 0; q0 < nstage; q0 += 4 * kGroupSlots) {
                int4 v4[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int q = min(q0 + k * kGroupSlots + tid, nstage - 1);
                    const int4 ce = s_cell[s_seg[q]];
                    v4[k] = gip[ce.x + q - ce.z];
                }
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int q = q0 + k * kGroupSlots + tid;
                    if (q < nstage) s_pj[q] = make_uint2(coarse_xy(v4[k]), coarse_z(v4[k]));
                }
            }
            __syncthreads();
        }
        int cnt = 0, qself = 0, pend = -1;
        if (valid && staged) {
            {  // the atom's own staged index (its own cell is window cell win_of(13))
                const int us = s_wmap[win_of(13)];
                if (us < nu) qself = s_cell[us].z + (i - s_cell[us].x);
            }
            // exclusions as staged indices: the atom itself and its first four
            // bonded partners (none, or a partner outside the 27 cells, which is
            // no candidate: INT_MAX)
            int xs[5] = {qself, INT_MAX, INT_MAX, INT_MAX, INT_MAX};
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int q = ex[b] >= 0 ? partner_q(ex[b]) : -1;
                xs[b + 1] = q >= 0 ? q : INT_MAX;
            }
            int nxu = 1;  // entries from nxu on are INT_MAX in every lane of the wave
#pragma unroll
            for (int b = 1; b < 5; b++) nxu = __ballot(xs[b] != INT_MAX) ? b + 1 : nxu;
            auto slot_of = [&](int q) __attribute__((always_inline)) {
                const int4 ce = s_cell[s_seg[q]];
                return ce.x + q - ce.z;
            };
            // Lower bounds (coarse units) of the distance along y and z from
            // this atom to any atom of the cell below and above its own: its
            // offset from its cell's bounds floor(c 2^16 / nx), less 2 units
            // (the truncation of the atom's coordinate and of the bound, and
            // the rounding of the float cell coordinate).  With 5 or more cells
            // per axis these offsets, taken as signed 16-bit differences, do
            // not wrap.
            int yb[2] = {0, 0}, zb[2] = {0, 0};
            const bool trim = nx >= 5;
            if (trim) {
                auto cb = [&](int c) { return (int)(((unsigned)c << 16) / (unsigned)nx); };
                auto off16 = [](int a, int b) { return max(0, (int)(short)(unsigned short)(a - b) - 2); };
                const int yc = (int)(cxy >> 16), zc = (int)cz16;
                yb[0] = off16(yc, cb(cy)); yb[1] = off16(cb(cy + 1), yc);
                zb[0] = off16(zc, cb(cz)); zb[1] = off16(cb(cz + 1), zc);
            }
            // x offset of staged atom q from this atom (coarse units, signed)
            auto dxc = [&](int q) { return (int)(short)(unsigned short)(s_pj[q].x - cxy); };
            // first q in [lo, hi) with dxc(q) > v: a cell's atoms are in
            // ascending x, and so are the three cells of a row in x order
            auto first_above = [&](int lo, int hi, int v) {
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (dxc(mid) <= v) lo = mid + 1;
                    else hi = mid;
                }
                return lo;
            };
            auto scan_rows = [&](auto w16) __attribute__((always_inline)) {
                constexpr bool W16 = decltype(w16)::value;
                auto scan_range = [&](int z, int e) __attribute__((always_inline)) {
                    for (int q0 = z; q0 < e; q0 += kBuildChunk) {
                        // a whole chunk from LDS (entries past the range are
                        // masked out, and the array is padded by a chunk)
                        uint2 jp[kBuildChunk];
#pragma unroll
                        for (int k = 0; k < kBuildChunk; k++) jp[k] = s_pj[q0 + k];
                        // bit xs[b] - q0 for each exclusion; one outside the
                        // chunk lands on a bit from 8 to 31, which is dropped
                        auto xbit = [&](int v) { return 1u << min((unsigned)(v - q0), 31u); };
                        unsigned x = xbit(xs[0]) | xbit(xs[1]) | xbit(xs[2]);
                        if (__builtin_expect(nxu > 3, 0)) x |= xbit(xs[3]) | xbit(xs[4]);
                        for (int b = 4; b < nb && b < d.bcap; b++) {  // partners past the first four
                            const unsigned o = (unsigned)(partner_q(bt[b * kTile] & kSlotMask) - q0);
                            x |= (o < (unsigned)kBuildChunk) ? (1u << o) : 0u;
                        }
                        build_chunk_q<W16>(jp, cxy, cz16, clo, (unsigned)(chi - clo), gpos, q0, e - q0, pi, L, invL,
                                           rcs2, x, out, out16, kcap, cnt, s_used, pend, slot_of);
                    }
                };
                // the 27 cells in (dz, dy, dx) order; the three cells of a
                // (dz, dy) row are consecutive union cells (hence one contiguous
                // staged range) unless the window wraps that axis
                #pragma unroll 1
                for (int r = 0; r < 9; r++) {
                    const int ua = s_wmap[win_of(3 * r)], ub = s_wmap[win_of(3 * r + 1)],
                              uc = s_wmap[win_of(3 * r + 2)];
                    if (ua >= nu || ub >= nu || uc >= nu) {  // every neighbor cell was marked; guard regardless
                        err |= ERR_INTERNAL;
                        continue;
                    }
                    const int4 A = s_cell[ua], B = s_cell[ub], Cc = s_cell[uc];
                    const bool one = ub == ua + 1 && uc == ub + 1;
                    // Only atoms within xr of this atom along x can be in the
                    // list: the distance along y and z to the row's cells leaves
                    // at most sqrt(rb2 - dy^2 - dz^2) along x, and the coarse x
                    // difference is within one unit of the true one.  A row with
                    // nothing left is skipped whole.
                    int xr = 0;
                    if (trim) {
                        const int dyb = (r % 3 == 1) ? 0 : yb[r % 3 >> 1], dzb = (r / 3 == 1) ? 0 : zb[r / 3 >> 1];
                        const int rem = rb2 - dyb * dyb - dzb * dzb;
                        if (rem < 0) continue;
                        xr = (int)ceilf(sqrtf((float)rem)) + 2;
                    }
                    if (one) {
                        int z = A.z, e = Cc.z + Cc.y;
                        if (trim) {
                            z = first_above(z, e, -xr - 1);
                            e = first_above(z, e, xr);
                        }
                        scan_range(z, e);
                    } else {
                        #pragma unroll 1
                        for (int t = 0; t < 3; t++) {
                            const int4 ce = t == 0 ? A : (t == 1 ? B : Cc);
                            int z = ce.z, e = ce.z + ce.y;
                            if (trim) {
                                z = first_above(z, e, -xr - 1);
                                e = first_above(z, e, xr);
                            }
                            scan_range(z, e);
                        }
                    }
                }
            };
            if (d.idx16) scan_rows(BoolC<true>{});
            else scan_rows(BoolC<false>{});
            // the deferred candidate: one load for the wave; outside -> the atom's
            // own entry (its staged index), which the force pass and the pair
            // readout skip
            if (pend >= 0) {
                const int pos = pend >> 12, q = pend & 0xFFF;
                const bool in = dist2_ref(pi, gpos[slot_of(q)], L, invL) < rcs2;
                if (!in && pos < kcap) {
                    if (d.idx16) out16[pos * kTile] = (unsigned short)qself;
                    else out[pos * kTile] = qself;
                }
            }
            if (qself < 0 || qself >= nstage) {
                err |= ERR_INTERNAL;
                qself = 0;
            }
        } else if (valid) {
            // global-memory scan with the same visiting order
            auto key_of = [&](int c) {
                return (int)cell_key(wrap_cell(cx, c % 3 - 1, nx), wrap_cell(cy, (c / 3) % 3 - 1, nx),
                                     wrap_cell(cz, c / 9 - 1, nx));
            };
            for (int c = 0; c < 27; c++) {
                const int k = key_of(c);
                bool seen = false;
                if (nx < 3)
                    for (int v = 0; v < c; v++) seen |= (key_of(v) == k);
                if (seen) continue;
                const int s0 = cstart[k], e0 = cstart[k + 1];
                for (int q0 = s0; q0 < e0; q0 += kBuildChunk) {
                    float4 pj[kBuildChunk];
#pragma unroll
                    for (int k2 = 0; k2 < kBuildChunk; k2++) pj[k2] = gpos[min(q0 + k2, e0 - 1)];
                    if (d.idx16)
                        build_chunk<true>(pj, q0, e0 - q0, i, pi, L, invL, rcs2, ex, nb, d.bcap, bt, out, out16,
                                          kcap, cnt, q0, nullptr);
                    else
                        build_chunk<false>(pj, q0, e0 - q0, i, pi, L, invL, rcs2, ex, nb, d.bcap, bt, out, out16,
                                           kcap, cnt, q0, nullptr);
                }
            }
        }
        const int stored = min(cnt, kcap);
        const int padded = valid ? ((stored + kLJTail - 1) / kLJTail) * kLJTail : 0;
        int selfent = i;  // the padding entry: the atom itself (rejected by r^2 = 0)
        bool compact = false;
        if (staged) {
            // the group's neighbor set: used staged atoms, every atom itself and
            // its bonded partners (the force kernel reads them from the set too),
            // numbered in staged (slot) order
            int pmiss = 0;
            if (valid) {
                s_used[qself] = 1;
                const int hv = gview(d.bhdr)[i], nbk = hv & 0xFF, nak = (hv >> 8) & 0xFF;
                const FMD_GLOBAL int2* at = gview(d.ang) + (size_t)t.lt * d.acap * kTile + lane;
                for (int b = 0; b < nbk; b++) {
                    const int q = partner_q(bt[b * kTile] & kSlotMask);
                    if (q < 0) pmiss = 1;
                    else s_used[q] = 1;
                }
                for (int a = 0; a < nak; a++) {
                    const int2 e = at[a * kTile];
                    const int q1 = partner_q(e.x & kSlotMask), q2 = partner_q(e.y & kSlotMask);
                    if (q1 < 0 || q2 < 0) pmiss = 1;
                    if (q1 >= 0) s_used[q1] = 1;
                    if (q2 >= 0) s_used[q2] = 1;
                }
            }
            // the scan's last reads of s_pj precede the s_cidx writes below
            pmiss = __syncthreads_or(pmiss);
            const int per = (nstage + kGroupSlots - 1) / kGroupSlots;
            const int q0 = min(tid * per, nstage), q1 = min(q0 + per, nstage);
            int c = 0;
            for (int q = q0; q < q1; q++) c += s_used[q];
            int ncomp = 0;
            int idx = block_excl_scan(c, s_wsum, &ncomp);
            compact = ncomp <= kGroupStage && !d.nostage && !pmiss;
            for (int q = q0; q < q1; q++)
                if (s_used[q]) {
                    s_cidx[q] = (unsigned short)idx;
                    if (compact) {
                        const int4 ce = s_cell[s_seg[q]];
                        const int slot = ce.x + q - ce.z;
                        const size_t at = (size_t)grp * kGroupStage + idx;
                        if (d.idx16) ((FMD_GLOBAL unsigned short*)gmut(d.gstage))[at] = (unsigned short)slot;
                        else gmut(d.gstage)[at] = slot;
                    }
                    idx++;
                }
            if (tid == 0) gmut(d.gcount)[grp] = compact ? ncomp : -1;
        } else if (tid == 0) {
            gmut(d.gcount)[grp] = -1;
        }
        // Placement: the group's rows go to columns in descending order of
        // their padded length (ties by slot), so the 64 rows a wave streams
        // are of nearly equal length; gperm[column] names the atom.  Rows are
        // moved a band of kPlaceBand entries at a time (every read of a band
        // precedes every write), mapping staged indices to LDS byte offsets
        // (or to slots when the neighbor set did not fit).
        s_len[tid] = padded;
        __syncthreads();
        if (valid && compact) {
            selfent = (int)s_cidx[qself] << 4;
            const int hv = gview(d.bhdr)[i], nbk = hv & 0xFF, nak = (hv >> 8) & 0xFF;
            const FMD_GLOBAL int2* at = gview(d.ang) + (size_t)t.lt * d.acap * kTile + lane;
            // bonded entries of a staged group: LDS byte offsets of the partners
            // in the set instead of slots (role and type bits kept)
            for (int b = 0; b < nbk; b++) {
                const int e = bt[b * kTile];
                gmut(d.bnd)[((size_t)t.lt * d.bcap + b) * kTile + lane] =
                    ((int)s_cidx[partner_q(e & kSlotMask)] << 4) | (e & ~kSlotMask);
            }
            for (int a = 0; a < nak; a++) {
                const int2 e = at[a * kTile];
                gmut(d.ang)[((size_t)t.lt * d.acap + a) * kTile + lane] =
                    make_int2(((int)s_cidx[partner_q(e.x & kSlotMask)] << 4) | (e.x & ~kSlotMask),
                              ((int)s_cidx[partner_q(e.y & kSlotMask)] << 4) | (e.y & ~kSlotMask));
            }
        }
        int rank = 0, maxpad = 0;
        for (int b = 0; b < kGroupSlots; b++) {
            const int lb = s_len[b];
            rank += (lb > padded || (lb == padded && b < tid)) ? 1 : 0;
            maxpad = max(maxpad, lb);
        }
        gmut(d.gperm)[(size_t)grp * kGroupSlots + rank] = (unsigned char)tid;
        // every row of a wave is padded to the wave's longest (its first
        // column's), whose length the force pass reads as one uniform bound
        __syncthreads();  // every read of s_len above precedes the writes below
        s_len[rank] = padded;
        __syncthreads();
        const int wpad = s_len[rank & ~(kTile - 1)];
        if ((rank & (kTile - 1)) == 0) gmut(d.tlen)[grp * kGroupTiles + rank / kTile] = wpad;
        auto entry = [&](int q) {  // a staged index or slot as the force kernel reads it
            if (!staged) return q;
            if (compact) return (int)s_cidx[q] << 4;
            return s_cell[s_seg[q]].x + q - s_cell[s_seg[q]].z;
        };
        // 32-bit rows, in place: every read of a band precedes every write
        auto place = [&](auto* src, auto* dst) {
            for (int k0 = 0; k0 < maxpad; k0 += kPlaceBand) {
                int v[kPlaceBand];
#pragma unroll
                for (int e = 0; e < kPlaceBand; e++) {
                    const int k = k0 + e;
                    v[e] = (k < stored) ? entry((int)src[k * kTile]) : selfent;
                }
                __syncthreads();
#pragma unroll
                for (int e = 0; e < kPlaceBand; e++)
                    if (k0 + e < wpad) dst[(k0 + e) * kTile] = v[e];
                __syncthreads();
            }
        };
        // destination column: the same addressing as out / out16 (tile base in int units)
        FMD_GLOBAL int* dtile = gmut(d.nbr) + (size_t)(grp * kGroupTiles + rank / kTile) * d.kcap * kTile;
        if (d.idx16) {
            // from this lane's row in the upper half to its column in the lower
            // half: no entry is both read and written, so no barriers
            // (the next band's loads are in flight while a band is converted)
            FMD_GLOBAL unsigned short* dst = (FMD_GLOBAL unsigned short*)dtile + (rank % kTile);
            int cur[kPlaceBand], nxt[kPlaceBand];
            auto load = [&](int k0, int (&to)[kPlaceBand]) {
#pragma unroll
                for (int e = 0; e < kPlaceBand; e++) to[e] = (k0 + e < stored) ? (int)out16[(k0 + e) * kTile] : 0;
            };
            load(0, cur);
            for (int k0 = 0; k0 < wpad; k0 += kPlaceBand) {
                if (k0 + kPlaceBand < wpad) load(k0 + kPlaceBand, nxt);
#pragma unroll
                for (int e = 0; e < kPlaceBand; e++) {
                    const int k = k0 + e;
                    if (k < wpad) dst[k * kTile] = (unsigned short)((k < stored) ? entry(cur[e]) : selfent);
                }
#pragma unroll
                for (int e = 0; e < kPlaceBand; e++) cur[e] = nxt[e];
            }
        } else {
            place(out, dtile + (rank % kTile));
        }
        if (valid) {
            if (cnt > kcap) err |= ERR_NBR_OVERFLOW;
            gmut(d.nnbr)[i] = stored;
        }
        err = wave_or(err);
        if (err && lane == 0) atomicOr(&st->err, err);
        __syncthreads();  // the next group reuses the LDS
    }
}

// LJ over this atom's Verlet rows (src/force/lj.cu:27-63), in grid units.
// With a = (float)(int)(Xi - Xj) the fixed-point displacement (grid step
// u = L * 2^-32), q = a.a - lo_g is (r^2 - 1e-10) / u^2: the cutoff test
// compares q with thresholds scaled by 1/u^2 (grid_s, grid_inside),
// 1/q = u^2 / r^2 (to a relative 1e-10 / r^2) feeds the LJ terms with
// sigma^2 / u^2 and 24 eps / u^2 folded in, and the force and virial sums are
// kept in grid units and scaled by u and u^2 once per atom.  Every row of a
// wave is padded with the atom itself (q = 0 is rejected) to the wave's
// longest row, a multiple of kLJTail (tlen), so the wave runs whole chunks of
// kLJChunk and at most one of kLJTail, loaded unconditionally, under one
// uniform bound; the next chunk's entries are in flight while the current
// chunk's positions are read.  The cutoff masks 1/q,
// so there is no per-pair branch.  OBS adds the per-atom PE and virial (thermo
// steps and NPT), CAP the push-off pair-force cap, ONE_TYPE a single
// (eps, sigma) pair with eps and sigma nonzero.  With LDS the entries are byte
// offsets into the group's neighbor set staged in LDS (sp); otherwise they are
// slots, read with indexed buffer loads (the hardware scales the slot by the
// 16-byte record).
struct LJSum {
    float fx, fy, fz, pe, w0, w1, w2, w3, w4, w5;  // force, PE and virial of one atom's rows
};

template <bool OBS, bool CAP, bool ONE_TYPE, bool W16, bool LDS>
__device__ __forceinline__ LJSum lj_loop(const RepDesc& d, int lt, int lane, int i, int4 pi, float L,
                                         const int4* sp) {
    const int nw = __builtin_amdgcn_readfirstlane(d.tlen[lt]);  // uniform: the loop control is scalar
    // the tile's rows as a raw buffer: entry k of this lane sits at byte
    // (k * kTile + lane) * entry size; the row offset goes in the scalar
    // offset, so fetching the next chunk costs no vector address arithmetic
    const __amdgpu_buffer_rsrc_t nrs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(d.nbr + (size_t)lt * d.kcap * kTile), (short)0, d.kcap * kTile * (int)sizeof(int), kBufferWord3);
    const unsigned vlane = (unsigned)lane * (W16 ? 2u : 4u);
    auto entry = [&](int k) -> unsigned {
        if constexpr (W16) return (unsigned)__builtin_amdgcn_raw_buffer_load_b16(nrs, vlane, k * kTile * 2, 0);
        else return __builtin_amdgcn_raw_buffer_load_b32(nrs, vlane, k * kTile * 4, 0);
    };
    const __amdgpu_buffer_rsrc_t prs =
        __builtin_amdgcn_make_buffer_rsrc((void*)d.ipos, (short)sizeof(int4), d.nslots, kBufferWord3);
    const float u = L * 2.3283064365386963e-10f;
    const GridCut gc = grid_cut(d.st);
    const float iu2 = d.st->giu2;
    const float2* ljrow = d.lj + pi.w * d.ntypes;
    const float2 p0 = ONE_TYPE ? d.lj[0] : ljrow[0];
    const float s2g0 = p0.y * p0.y * iu2, c24g0 = 24.0f * p0.x * iu2, c2e0 = 2.0f * p0.x;
    const float capg = CAP ? d.fcap / u : 0.0f;  // fcap / r = (fcap / u) / sqrt(q)
    // one (eps, sigma): force / r = K sr2 sr6 (2 sr6 - 1) with K = 24 eps / sigma^2,
    // summed without K (applied once per atom); the +-1000 clamp becomes +-1000 / K
    constexpr bool fold = ONE_TYPE && !CAP;  // ONE_TYPE only with eps, sigma != 0
    const float kf = fold ? d.lj_kf : 1.0f;
    const float fclamp = fold ? d.lj_fclamp : 1000.0f;
    float gx = 0.f, gy = 0.f, gz = 0.f, pe = 0.f;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;
    auto pair = [&](const uint4 pj) __attribute__((always_inline)) {
        const float ax = grid_d(pi.x, pj.x), ay = grid_d(pi.y, pj.y), az = grid_d(pi.z, pj.z);
        const float q = grid_s(ax, ay, az, gc.lo_g);
        // forces only, one (eps, sigma): q == 0 may count as inside (its
        // force/r is the clamp and finite), one compare fewer per pair
        const bool in = (fold && !OBS) ? grid_inside0(q, gc.span) : grid_inside(q, gc.span);
        float s2g = s2g0, c24g = c24g0, c2e = c2e0;
        if (!ONE_TYPE) {
            const float2 p = ljrow[pj.w];
            s2g = p.y * p.y * iu2; c24g = 24.0f * p.x * iu2; c2e = 2.0f * p.x;
        }
        const float qinv = in ? __builtin_amdgcn_rcpf(q) : 0.0f;
        const float sr2 = s2g * qinv;
        const float sr6 = sr2 * sr2 * sr2;
        const float sr12 = sr6 * sr6;
        float fr;
        if constexpr (fold) fr = (sr2 * sr6) * fmaf(sr6, 2.0f, -1.0f);  // force / r over K
        else fr = c24g * qinv * (2.0f * sr12 - sr6);        // force / r in real units
        fr = __builtin_amdgcn_fmed3f(fr, -fclamp, fclamp);  // src/force/lj.cu:48
        if (CAP) {
            const float cap = capg * __builtin_amdgcn_rsqf(q);
            fr = fminf(fmaxf(fr, -cap), cap);
        }
        gx += fr * ax; gy += fr * ay; gz += fr * az;
        // keep the type word live so the neighbor read stays 16 bytes:
        // gfx950 serves randomly addressed 12-byte LDS reads at about half
        // the rate of 16-byte ones (the empty asm emits no instruction)
        if constexpr (LDS && ONE_TYPE) asm volatile("" ::"v"(pj.w));
        if (OBS) {
            const float hx = fr * ax, hy = fr * ay, hz = fr * az;
            pe += c2e * (sr12 - sr6);
            v0 += hx * ax; v1 += hx * ay; v2 += hx * az;
            v3 += hy * ay; v4 += hy * az; v5 += hz * az;
        }
    };
    auto neighbor = [&](unsigned e) __attribute__((always_inline)) -> uint4 {
        if constexpr (LDS) {  // the row holds the LDS byte offset of the staged neighbor
            const int4 v = *(const int4*)((const char*)sp + e);
            return make_uint4((unsigned)v.x, (unsigned)v.y, (unsigned)v.z, (unsigned)v.w);
        } else {
            const u32x4 v = struct_load_b128(prs, e, 0, 0, 0);
            return make_uint4(v[0], v[1], v[2], v[3]);
        }
    };
    // nw, the wave's padded row length (a multiple of kLJTail, the same for
    // every lane: the list build pads each row of a wave to it), bounds whole
    // chunks of kLJChunk and a last chunk of kLJTail; entries fetched past nw
    // (within the row area, or zero from the bounded buffer) are never used
    unsigned nxt[kLJChunk];
    if (nw > 0) {
#pragma unroll
        for (int k = 0; k < kLJChunk; k++) nxt[k] = entry(k);
    }
    int k0 = 0;
    for (; k0 + kLJChunk <= nw; k0 += kLJChunk) {
        uint4 pjs[kLJChunk];
#pragma unroll
        for (int k = 0; k < kLJChunk; k++) pjs[k] = neighbor(nxt[k]);
        if (k0 + kLJChunk < nw) {
#pragma unroll
            for (int k = 0; k < kLJChunk; k++) nxt[k] = entry(k0 + kLJChunk + k);
        }
#pragma unroll
        for (int k = 0; k < kLJChunk; k++) pair(pjs[k]);
    }
    if (k0 < nw) {  // nw - k0 == kLJTail
        uint4 pjs[kLJTail];
#pragma unroll
        for (int k = 0; k < kLJTail; k++) pjs[k] = neighbor(nxt[k]);
#pragma unroll
        for (int k = 0; k < kLJTail; k++) pair(pjs[k]);
    }
    const float fu = kf * u, hu2 = 0.5f * kf * u * u;
    return LJSum{gx * fu, gy * fu, gz * fu, pe, v0 * hu2, v1 * hu2, v2 * hu2, v3 * hu2, v4 * hu2, v5 * hu2};
}

#define LJ_ARGS d, t.lt, lane, i, pi, L, sp
template <bool W16, bool LDS>
__device__ __forceinline__ LJSum lj_variants(const RepDesc& d, const Tile& t, int lane, int i, int4 pi,
                                             float L, bool obs, const int4* sp) {
    const bool cap = d.fcap > 0.0f, one = d.ntypes == 1 && d.lj[0].x != 0.0f && d.lj[0].y != 0.0f;
    if (!cap && one)
        return obs ? lj_loop<true, false, true, W16, LDS>(LJ_ARGS) : lj_loop<false, false, true, W16, LDS>(LJ_ARGS);
    if (!cap)
        return obs ? lj_loop<true, false, false, W16, LDS>(LJ_ARGS)
                   : lj_loop<false, false, false, W16, LDS>(LJ_ARGS);
    return obs ? lj_loop<true, true, false, W16, LDS>(LJ_ARGS) : lj_loop<false, true, false, W16, LDS>(LJ_ARGS);
}
#undef LJ_ARGS

// Tabulated pairs over the same Verlet rows (src/force/table.cu:58-118).  A
// pair inside the cutoff (the grid-unit test of the LJ loop) whose types have a
// table gets the linear interpolant of force / r and energy at r clamped to
// [rmin, rmax]: node k sits at rmin + k dr, k = (int)((r - rmin) / dr) clamped
// to [0, n - 2], weight t = (r - (rmin + k dr)) / dr.  Half the pair energy and
// half the pair virial go to this atom, as in the LJ loop.
template <bool W16>
__device__ __forceinline__ LJSum table_loop(const RepDesc& d, int lt, int lane, int i, int4 pi, float L, const int4* sp) {
    const int nn = d.nnbr[i];
    const int* nl = d.nbr + (size_t)lt * d.kcap * kTile + lane;
    const unsigned short* nl16 = (const unsigned short*)(d.nbr + (size_t)lt * d.kcap * kTile) + lane;
    const float u = L * 2.3283064365386963e-10f;
    const GridCut gc = grid_cut(d.st);
    const int* trow = d.tab_idx + pi.w * d.ntypes;
    LJSum s{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < nn; k++) {
        const int j = W16 ? (int)nl16[k * kTile] : nl[k * kTile];
        const int4 pj = sp ? sp[j >> 4] : d.ipos[j];  // LDS byte offset or slot
        const int tb = trow[pj.w];
        if (tb < 0) continue;
        const float ax = grid_d(pi.x, pj.x), ay = grid_d(pi.y, pj.y), az = grid_d(pi.z, pj.z);
        if (!grid_inside(grid_s(ax, ay, az, gc.lo_g), gc.span)) continue;
        const float dx = ax * u, dy = ay * u, dz = az * u;
        const float4 tp = d.tab_parm[tb];
        const int2 tn = d.tab_no[tb];
        const float r = fminf(fmaxf(sqrtf(sq3(dx, dy, dz)), tp.x), tp.y);
        const int idx = min(max((int)((r - tp.x) * tp.w), 0), tn.x - 2);
        const float t = (r - (tp.x + (float)idx * tp.z)) * tp.w;
        const float2 p0 = d.tab_fe[tn.y + idx], p1 = d.tab_fe[tn.y + idx + 1];
        const float fs = p0.x + t * (p1.x - p0.x);
        const float e = p0.y + t * (p1.y - p0.y);
        const float hx = 0.5f * fs * dx, hy = 0.5f * fs * dy, hz = 0.5f * fs * dz;
        s.fx += fs * dx; s.fy += fs * dy; s.fz += fs * dz;
        s.pe += 0.5f * e;
        s.w0 += hx * dx; s.w1 += hx * dy; s.w2 += hx * dz;
        s.w3 += hy * dy; s.w4 += hy * dz; s.w5 += hz * dz;
    }
    return s;
}

// atan2(y, x) for y >= 0, the angle between two bond vectors in [0, pi], by
// the range reduction and polynomial of the Cephes atanf (S. L. Moshier): the
// ratio r = min(|x|, y) / max(|x|, y) in [0, 1] becomes t = (r - 1) / (r + 1)
// above tan(pi/8), atan(t) = t + t z P(z) with z = t^2, then the octant is
// undone.  Within 4 ulp of an FP64 atan2 on random pairs; near a straight
// angle (small y, x < 0) within one ulp of pi.
__device__ __forceinline__ float angle_atan2(float y, float x) {
    const float ax = fabsf(x);
    const float mn = fminf(ax, y), mx = fmaxf(ax, y);
    const float r = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
    const bool big = r > 0.41421356237f;  // tan(pi/8)
    const float t = big ? (r - 1.0f) * __builtin_amdgcn_rcpf(r + 1.0f) : r;
    const float z = t * t;
    const float p = fmaf(fmaf(fmaf(8.05374449538e-2f, z, -1.38776856032e-1f), z, 1.99777106478e-1f), z,
                         -3.33329491539e-1f);
    float a = fmaf(p * z, t, t);
    if (big) a += 0.78539816339744831f;     // pi/4
    if (y > ax) a = 1.5707963267948966f - a;  // pi/2
    if (x < 0.0f) a = 3.1415926535897931f - a;
    return a;
}

// ---------------------------------------------------------------------------
// Force kernel: LJ over the Verlet list (src/force/lj.cu:27-63), FENE+WCA
// (src/force/fene.cu:21-65) and harmonic angle (src/force/angle.cu:22-79) as
// per-atom gathers (no float atomics), then the post-force half kick
// (src/integrate/langevin.cu:109-130).  Writes force[i] = (F, pe_i) and 13
// FP64 partials per wave: pe, virial[6], kinetic tensor[6] (after the kick).
// ---------------------------------------------------------------------------
template <int TERMS, bool LIST = false>
__global__ void __launch_bounds__(kMaxBlockK) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_force(const RepDesc* __restrict__ descs, const Tile* __restrict__ tiles, int t0, int ntl, int kick, int mode,
        int sel, const int* __restrict__ rlist, const int* __restrict__ nrt) {
    constexpr int terms = TERMS;
    __shared__ int4 s_pos[(TERMS & 1) ? kGroupStage : 1];  // the group's neighbor set
    // One tile per wave.  t0 and ntl are whole groups and a workgroup (1, 2 or
    // 4 waves) lies in one group, so the returns below are uniform over the
    // workgroup.
    auto tile = [&](int w) __attribute__((always_inline)) {
        const Tile t = tiles[w];
        const RepDesc& d = descs[t.rep];
        RepState* st = d.st;
        if (mode == FM_STEP && !st->active) return;
        if (sel == FS_STAYING && st->rebuild) return;
        const int lane = lane_id();
        // this lane computes the atom whose row the list build placed in its column
        const int grp = t.lt / kGroupTiles;
        const int i = grp * kGroupSlots + d.gperm[t.lt * kTile + lane];
        // stage the group's neighbor set (coalesced slot list, one 16-byte position each)
        const int nst = (terms & 1) ? d.gcount[grp] : -1;
        if (nst >= 0) {
            const int nth = blockDim.x, last = nst - 1;
            const FMD_GLOBAL int4* gp = gview(d.ipos);
            auto stage = [&](auto gs) {  // four slots, then four positions, in flight at once
                for (int q = threadIdx.x; q < nst; q += 4 * nth) {
                    const int qb = q + nth, qc = q + 2 * nth, qd = q + 3 * nth;
                    const int sa = gs[q], sb = gs[min(qb, last)], sc = gs[min(qc, last)], sd = gs[min(qd, last)];
                    const int4 va = gp[sa], vb = gp[sb], vc = gp[sc], vd = gp[sd];
                    s_pos[q] = va;
                    if (qb < nst) s_pos[qb] = vb;
                    if (qc < nst) s_pos[qc] = vc;
                    if (qd < nst) s_pos[qd] = vd;
                }
            };
            if (d.idx16) stage((const FMD_GLOBAL unsigned short*)gview(d.gstage) + (size_t)grp * kGroupStage);
            else stage(gview(d.gstage) + (size_t)grp * kGroupStage);
            __syncthreads();
        }
        const int4* sp = nst >= 0 ? s_pos : nullptr;
        float fx = 0.f, fy = 0.f, fz = 0.f, pe = 0.f;
        float w0 = 0.f, w1 = 0.f, w2 = 0.f, w3 = 0.f, w4 = 0.f, w5 = 0.f;
        // Langevin steps that record nothing need none of the FP64 partials (no
        // thermostat reads the kinetic energy; the host reevaluates unobserved
        // steps on request), and neither does a forces-only launch
        const bool partials = mode == FM_OBS || (mode == FM_STEP && (d.ensemble != 0 || st->want_obs));
        double k0 = 0, k1 = 0, k2 = 0, k3 = 0, k4 = 0, k5 = 0;
        double vkey = 0.0, fkey = 0.0;  // max_key of |v|^2 and |F|^2 (src/analysis/thermo.cu:26-75)
        int err = 0;
        if (i < d.natoms) {
            const int4 pi = d.ipos[i];
            const int ti = pi.w;
            const float L = st->Lp;
            const float ulen = L * 2.3283064365386963e-10f;  // fixed-point step, L * 2^-32
            const bool obs = (mode == FM_OBS) || (mode == FM_STEP && st->want_obs);
            if (terms & 1) {
                LJSum a;
                if (sp) {  // the array itself: its LDS address is a constant of the kernel
                    if (d.idx16) a = lj_variants<true, true>(d, t, lane, i, pi, L, obs, s_pos);
                    else a = lj_variants<false, true>(d, t, lane, i, pi, L, obs, s_pos);
                } else {
                    if (d.idx16) a = lj_variants<true, false>(d, t, lane, i, pi, L, obs, sp);
                    else a = lj_variants<false, false>(d, t, lane, i, pi, L, obs, sp);
                }
                fx = a.fx; fy = a.fy; fz = a.fz; pe = a.pe;
                w0 = a.w0; w1 = a.w1; w2 = a.w2; w3 = a.w3; w4 = a.w4; w5 = a.w5;
                if (d.ntables) {
                    const LJSum b = d.idx16 ? table_loop<true>(d, t.lt, lane, i, pi, L, sp)
                                            : table_loop<false>(d, t.lt, lane, i, pi, L, sp);
                    fx += b.fx; fy += b.fy; fz += b.fz; pe += b.pe;
                    w0 += b.w0; w1 += b.w1; w2 += b.w2; w3 += b.w3; w4 += b.w4; w5 += b.w5;
                }
            }
            const int hdr = (terms & 6) ? d.bhdr[i] : 0;
            // a bonded entry of a staged group is the partner's LDS byte offset in
            // the group's neighbor set (read from the set, or through its slot
            // list when this variant did not stage it); otherwise a slot
            const bool bstaged = (terms & 6) && (sp || d.gcount[grp] >= 0);
            const FMD_GLOBAL unsigned short* gs16 =
                (const FMD_GLOBAL unsigned short*)gview(d.gstage) + (size_t)grp * kGroupStage;
            const FMD_GLOBAL int* gs32 = gview(d.gstage) + (size_t)grp * kGroupStage;
            auto partner = [&](int e) __attribute__((always_inline)) -> int4 {
                const int v = e & kSlotMask;
                if (sp) return *(const int4*)((const char*)s_pos + v);
                if (bstaged) return d.ipos[d.idx16 ? (int)gs16[v >> 4] : gs32[v >> 4]];
                return d.ipos[v];
            };
            // The bonded terms in the order bonds, then angles, each over its rows;
            // the entries and partner positions of the first kBondAhead bonds and
            // kAngleAhead angles (all of a Kremer-Grest bead's) are requested
            // together, so their latencies overlap instead of forming one chain per
            // term.
            const int nb = (terms & 2) ? (hdr & 0xFF) : 0;
            const int na = (terms & 4) ? ((hdr >> 8) & 0xFF) : 0;
            const int* bt = d.bnd + (size_t)(i / kTile) * d.bcap * kTile + (i % kTile);
            const int2* at = d.ang + (size_t)(i / kTile) * d.acap * kTile + (i % kTile);
            auto bond_term = [&](int e, int4 pj) __attribute__((always_inline)) {
                const float4 bp = d.bparm[(unsigned)e >> kSlotBits];
                const float dx = fix_diff(pi.x, pj.x, ulen);
                const float dy = fix_diff(pi.y, pj.y, ulen);
                const float dz = fix_diff(pi.z, pj.z, ulen);
                float r2 = sq3(dx, dy, dz);
                const float R02 = bp.y * bp.y;
                // a bond at or beyond R0 is a failure of the run, except during the
                // capped soft push-off, where the clamp below keeps it finite
                if (!(r2 < R02) && !(d.fcap > 0.0f)) err |= ERR_STRETCH;
                r2 = fminf(fmaxf(r2, 1e-6f), R02 * 0.9999f);  // src/force/fene.cu:33
                // divisions as hardware reciprocals (1 ulp), as the reference's
                // fast-math build does; PE and virial only where observed
                float ff = -bp.x * R02 * __builtin_amdgcn_rcpf(R02 - r2);
                const float sig2 = bp.w * bp.w;
                const bool wca = r2 < sig2 * 1.2599210498948732f;
                float sr6 = 0.0f, sr12 = 0.0f;
                if (wca) {
                    const float r2inv = __builtin_amdgcn_rcpf(r2);
                    const float sr2 = sig2 * r2inv;
                    sr6 = sr2 * sr2 * sr2;
                    sr12 = sr6 * sr6;
                    ff += 24.0f * bp.z * r2inv * (2.0f * sr12 - sr6);
                }
                fx += ff * dx; fy += ff * dy; fz += ff * dz;
                if (obs) {
                    float bpe = -0.5f * bp.x * R02 * logf(1.0f - r2 / R02);
                    if (wca) bpe += 4.0f * bp.z * (sr12 - sr6) + bp.z;
                    const float hx = 0.5f * ff * dx, hy = 0.5f * ff * dy, hz = 0.5f * ff * dz;
                    pe += 0.5f * bpe;
                    w0 += hx * dx; w1 += hx * dy; w2 += hx * dz;
                    w3 += hy * dy; w4 += hy * dz; w5 += hz * dz;
                }
            };
            auto angle_term = [&](int2 e, int4 q1, int4 q2) __attribute__((always_inline)) {
                const int role = (unsigned)e.x >> kSlotBits;
                const float2 ap = d.aparm[(unsigned)e.y >> kSlotBits];
                // angle atoms (A, B=center, C) from this atom's role
                // (assignments, not conditionals of two lvalues: a select of
                // addresses would keep pi in scratch memory)
                int4 pa = q1, pb = q1, pc = q2;
                if (role == 0) pa = pi;
                if (role == 1) pb = pi;
                else if (role == 2) pb = q2;
                if (role == 2) pc = pi;
                const float ax = fix_diff(pa.x, pb.x, ulen);
                const float ay = fix_diff(pa.y, pb.y, ulen);
                const float az = fix_diff(pa.z, pb.z, ulen);
                const float bx = fix_diff(pc.x, pb.x, ulen);
                const float by = fix_diff(pc.y, pb.y, ulen);
                const float bz = fix_diff(pc.z, pb.z, ulen);
                const float ra2 = ax * ax + ay * ay + az * az;
                const float rb2 = bx * bx + by * by + bz * bz;
                const float cxp = ay * bz - az * by, cyp = az * bx - ax * bz, czp = ax * by - ay * bx;
                const float cr = sqrtf(cxp * cxp + cyp * cyp + czp * czp);
                const float dot = ax * bx + ay * by + az * bz;
                const float theta = angle_atan2(cr, dot);
                const float inv_ab = rsqrtf(ra2) * rsqrtf(rb2);
                const float c = dot * inv_ab;
                const float s = fmaxf(cr * inv_ab, 1e-6f);
                const float dth = theta - ap.y;
                const float pref = 2.0f * ap.x * dth * __builtin_amdgcn_rcpf(s);
                const float ia2 = __builtin_amdgcn_rcpf(ra2), ib2 = __builtin_amdgcn_rcpf(rb2);
                const float fix = pref * (bx * inv_ab - c * ax * ia2);
                const float fiy = pref * (by * inv_ab - c * ay * ia2);
                const float fiz = pref * (bz * inv_ab - c * az * ia2);
                const float fkx = pref * (ax * inv_ab - c * bx * ib2);
                const float fky = pref * (ay * inv_ab - c * by * ib2);
                const float fkz = pref * (az * inv_ab - c * bz * ib2);
                if (role == 0) { fx += fix; fy += fiy; fz += fiz; }
                else if (role == 2) { fx += fkx; fy += fky; fz += fkz; }
                else { fx -= fix + fkx; fy -= fiy + fky; fz -= fiz + fkz; }
                if (obs) {
                    const float third = 1.0f / 3.0f;
                    pe += ap.x * dth * dth * third;
                    w0 += (fix * ax + fkx * bx) * third;
                    w1 += (fix * ay + fkx * by) * third;
                    w2 += (fix * az + fkx * bz) * third;
                    w3 += (fiy * ay + fky * by) * third;
                    w4 += (fiy * az + fky * bz) * third;
                    w5 += (fiz * az + fkz * bz) * third;
                }
            };
            constexpr int kBondAhead = 2, kAngleAhead = 3;
            const int eb0 = (nb > 0) ? bt[0] : 0, eb1 = (nb > 1) ? bt[kTile] : 0;
            const int2 ea0 = (na > 0) ? at[0] : make_int2(0, 0), ea1 = (na > 1) ? at[kTile] : make_int2(0, 0),
                       ea2 = (na > 2) ? at[2 * kTile] : make_int2(0, 0);
            int4 pb0 = pi, pb1 = pi, qa0 = pi, qc0 = pi, qa1 = pi, qc1 = pi, qa2 = pi, qc2 = pi;
            if (nb > 0) pb0 = partner(eb0);
            if (nb > 1) pb1 = partner(eb1);
            if (na > 0) { qa0 = partner(ea0.x); qc0 = partner(ea0.y); }
            if (na > 1) { qa1 = partner(ea1.x); qc1 = partner(ea1.y); }
            if (na > 2) { qa2 = partner(ea2.x); qc2 = partner(ea2.y); }
            if (nb > 0) bond_term(eb0, pb0);
            if (nb > 1) bond_term(eb1, pb1);
            for (int b = kBondAhead; b < nb; b++) {
                const int e = bt[b * kTile];
                bond_term(e, partner(e));
            }
            if (na > 0) angle_term(ea0, qa0, qc0);
            if (na > 1) angle_term(ea1, qa1, qc1);
            if (na > 2) angle_term(ea2, qa2, qc2);
            for (int a = kAngleAhead; a < na; a++) {
                const int2 e = at[a * kTile];
                angle_term(e, partner(e.x), partner(e.y));
            }
            // Langevin steps followed by another step of the run also give that
            // step's pre-force half kick (the same force, the same operations as
            // k_pre), so the force need not be stored and read back; the kinetic
            // partials are taken between the two kicks
            const bool ahead = kick && mode == FM_STEP && d.ensemble == 0 && st->step + 1 < st->stop;
            if (!ahead) d.force[i] = make_float4(fx, fy, fz, pe);
            float4 v = d.vel[i];
            const float m = d.mass[ti];
            const float hdtm = d.half_dt / m;
            if (kick) { v.x = fmaf(hdtm, fx, v.x); v.y = fmaf(hdtm, fy, v.y); v.z = fmaf(hdtm, fz, v.z); }
            if (partials) {
                const double md = m, vx = v.x, vy = v.y, vz = v.z;
                k0 = md * vx * vx; k1 = md * vx * vy; k2 = md * vx * vz;
                k3 = md * vy * vy; k4 = md * vy * vz; k5 = md * vz * vz;
                const int g = d.gid[i];
                vkey = max_key(v.x * v.x + v.y * v.y + v.z * v.z, g);
                fkey = max_key(fx * fx + fy * fy + fz * fz, g);
            }
            if (ahead) {  // bit for bit the kick of k_pre (explicit FMAs there too)
                v.x = fmaf(hdtm, fx, v.x); v.y = fmaf(hdtm, fy, v.y); v.z = fmaf(hdtm, fz, v.z);
                if (d.vlim > 0.0f) v = limit_speed(v, d.vlim);
            }
            if (kick) d.vel[i] = v;
        }
        if (partials) {
            double part[kNumSum] = {(double)pe, (double)w0, (double)w1, (double)w2, (double)w3, (double)w4,
                                    (double)w5, k0, k1, k2, k3, k4, k5};
#pragma unroll
            for (int c = 0; c < kNumSum; c++) part[c] = wave_sum_last(part[c]);
            vkey = wave_max_last(vkey);
            fkey = wave_max_last(fkey);
            if (lane == 63) {
#pragma unroll
                for (int c = 0; c < kNumSum; c++) d.wpart[(size_t)t.lt * kNumPart + c] = part[c];
                d.wpart[(size_t)t.lt * kNumPart + kNumSum] = vkey;
                d.wpart[(size_t)t.lt * kNumPart + kNumSum + 1] = fkey;
            }
        }
        if (err) atomicOr(&st->err, err);  // failures only
    };
    if constexpr (LIST) {
        // the listed tiles (the rebuild phases' list): each workgroup (four
        // waves: one group) takes listed groups in turn, the workgroups of one
        // XCD (a multiple of kXcds in the grid) from one contiguous eighth of
        // the list, so a replica's positions are gathered through one L2
        const int ng = *nrt / kGroupTiles, per = (ng + kXcds - 1) / kXcds, x = blockIdx.x % kXcds;
        const int g1 = min((x + 1) * per, ng), gs = gridDim.x / kXcds;
        for (int g = x * per + blockIdx.x / kXcds; g < g1; g += gs) {
            tile(rlist[g * kGroupTiles + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)]);
            __syncthreads();  // the next group's staging reuses s_pos
        }
    } else {
        // the grid's wave w takes tile t0 + w
        const int w = t0 + xcd_wave_in_grid();
        if (w < t0 + ntl) tile(w);
    }
}

// First reduction level: one workgroup sums the FP64 partials of kChunkTiles
// tiles of one replica in a fixed order (thread t takes tile t, then the DPP
// wave sum, then the four waves in order), so a large replica's sum is spread
// over many CUs.
__global__ void __launch_bounds__(256) k_chunk(const RepDesc* __restrict__ descs,
                                              const Tile* __restrict__ chunks, int nch, int mode) {
    constexpr int kN = kNumSum + 2;
    __shared__ double red[kN][4];
    const Tile ch = chunks[blockIdx.x];
    const RepDesc& d = descs[ch.rep];
    if (mode == BND_STEP && (!d.st->active || (d.ensemble == 0 && !d.st->want_obs))) return;  // nothing to sum
    const int tl = ch.lt * kChunkTiles + threadIdx.x;
    double a[kN];
#pragma unroll
    for (int c = 0; c < kN; c++) a[c] = 0.0;
    if (tl < d.ntiles)
#pragma unroll
        for (int c = 0; c < kN; c++) a[c] = d.wpart[(size_t)tl * kNumPart + c];
#pragma unroll
    for (int c = 0; c < kN; c++) a[c] = c < kNumSum ? wave_sum_last(a[c]) : wave_max_last(a[c]);
    if (lane_id() == 63)
#pragma unroll
        for (int c = 0; c < kN; c++) red[c][threadIdx.x >> 6] = a[c];
    __syncthreads();
    if (threadIdx.x < kN) {
        const int c = threadIdx.x;
        d.cpart[(size_t)ch.lt * kNumPart + c] =
            c < kNumSum ? ((red[c][0] + red[c][1]) + red[c][2]) + red[c][3]
                        : fmax(fmax(red[c][0], red[c][1]), fmax(red[c][2], red[c][3]));
    }
}

// Multiple-tau correlator push (src/analysis/correlator.cu:33-92), by the
// whole block: at each level thread (lag, c) adds sample * the sample lag
// pushes back (lags up to the ring's insert position, as the reference
// counts them), then thread 0 advances the ring and the block accumulator,
// which hands its mean to the next level every kCorrPoints samples.  smp is
// the sample (shared) and is replaced by the block mean on the way up.
__device__ void corr_push(CorrState* cs, double* smp) {
    __shared__ int cont;
    const int t = threadIdx.x, lag = t / 6, c = t % 6;
    for (int lv = 0; lv < kCorrLevels; lv++) {
        const int idx = cs->ins[lv];
        if (t < kCorrPoints * 6) {
            const double x = smp[c];
            const int base = lv * kCorrPoints;
            if (lag == 0) cs->buf[(base + idx) * 6 + c] = x;
            if (lag <= idx) {
                const double o = (lag == 0) ? x : cs->buf[(base + idx - lag) * 6 + c];
                cs->corr[(base + lag) * 6 + c] += x * o;
                if (c == 0) cs->counts[base + lag]++;
            }
        }
        __syncthreads();
        if (t == 0) {
            cs->ins[lv] = (idx + 1) % kCorrPoints;
            const int n = cs->bcount[lv] + 1;
            for (int k = 0; k < 6; k++) cs->acc[lv * 6 + k] += smp[k];
            cont = n >= kCorrPoints;
            if (cont) {
                for (int k = 0; k < 6; k++) {
                    smp[k] = cs->acc[lv * 6 + k] / (double)kCorrPoints;
                    cs->acc[lv * 6 + k] = 0.0;
                }
            }
            cs->bcount[lv] = cont ? 0 : n;
        }
        __syncthreads();
        if (!cont) break;
    }
}

__global__ void __launch_bounds__(128) k_corr_push(const RepDesc* __restrict__ d, Sample6 x) {
    __shared__ double smp[6];
    if (threadIdx.x < 6) smp[threadIdx.x] = x.v[threadIdx.x];
    __syncthreads();
    corr_push(d->corr, smp);
}

// The step's chain applications on a stream of kinetic energies, one lane,
// state carried from draw to draw (the device chain of nh_propagate_chain_kernel,
// src/integrate/nose_hoover.cu:86-178, as tests/test_nose_hoover.cu:343-439
// drives it).  carry = 0: one application per draw with KE ke[t]; carry = 1:
// two per draw, the post-force one with ke[t] and then the next pre-force one
// with the scaled kinetic energy ke[t] s^2 it leaves (chain_KE_carry).
__global__ void __launch_bounds__(64) k_nh_chain_run(RepState* st, const double* __restrict__ ke, int n,
                                                     int carry, double kT, int natoms, double half_dt,
                                                     double* __restrict__ scales) {
    if (threadIdx.x != 0) return;
    for (int t = 0; t < n; t++) {
        const double s = nh_chain(st, ke[t], kT, natoms, half_dt);
        if (carry) {
            scales[2 * t] = s;
            scales[2 * t + 1] = nh_chain(st, ke[t] * s * s, kT, natoms, half_dt);
        } else {
            scales[t] = s;
        }
    }
}

// Scalar tail of the per-replica boundary (thread 0): the post-force chain /
// barostat of this step, start-of-step scalars of the next, thermo record,
// step counter.  Returns 1 when this step's pressure tensor is a correlator
// sample (written to smp).
__device__ int boundary_tail(const RepDesc& d, RepState* st, int mode, const double (*red)[kBndThreads / 64],
                             double* smp) {
    double tot[kNumSum + 2];
    for (int c = 0; c < kNumSum + 2; c++) {
        tot[c] = 0.0;
        for (int w = 0; w < kBndThreads / 64; w++) tot[c] = c < kNumSum ? tot[c] + red[c][w] : fmax(tot[c], red[c][w]);
    }
    const double pe = tot[0];
    double W[6], K[6];
    for (int c = 0; c < 6; c++) { W[c] = tot[1 + c]; K[c] = tot[7 + c]; }
    const double KE = 0.5 * (K[0] + K[3] + K[5]);
    const bool reduced = !(mode == BND_STEP && d.ensemble == 0 && !st->want_obs);
    if (reduced) {
        st->vkey = tot[kNumSum];
        st->fkey = tot[kNumSum + 1];
    }
    if (mode == BND_REDUCE) {
        st->obs_step = st->step;
        st->pe = pe;
        for (int c = 0; c < 6; c++) { st->vir[c] = W[c]; st->kin[c] = K[c] * (double)st->v_pending * (double)st->v_pending; }
        return 0;
    }
    if (mode == BND_PRIME) {
        st->obs_step = st->step;
        st->pe = pe;
        for (int c = 0; c < 6; c++) { st->vir[c] = W[c]; st->kin[c] = K[c]; }
        st->ke_carry = KE;  // src/main.cu:373-386
        st->v_pending = 1.0f;
        st->rec_step = st->step;
        st->pre_ready = 0;
        return 0;
    }
    // BND_STEP: end of step cur
    const int64_t cur = st->step + 1;
    const double hdt = 0.5 * (double)d.dt, dt = (double)d.dt;
    double ks = 1.0;
    if (d.ensemble == 1) {
        const float frac = (float)(cur - 1) / (float)st->ramp_n;
        const float T = (float)st->T_start + ((float)st->T_stop - (float)st->T_start) * frac;
        const double s = nh_chain(st, KE, (double)T, d.natoms, hdt);
        st->ke_carry = KE * s * s;
        ks = s;
    } else if (d.ensemble == 2) {
        const float frac = (float)(cur - 1) / (float)st->ramp_n;
        const float T = (float)st->T_start + ((float)st->T_stop - (float)st->T_start) * frac;
        const float P = (float)st->P_start + ((float)st->P_stop - (float)st->P_start) * frac;
        const BaroOut o = npt_chain_baro(st, KE, W[0] + W[3] + W[5], true, (double)T, (double)P,
                                         d.natoms, hdt, dt);
        st->ke_carry = o.KE_scaled * o.nh_scale * o.nh_scale;
        ks = o.baro_scale * o.nh_scale;
        st->L = (float)o.L;
        st->invL = 1.0f / st->L;
    }
    st->v_pending = (float)ks;
    const double ks2 = (double)st->v_pending * (double)st->v_pending;
    if (d.ensemble != 0 || st->want_obs)  // else this step reduced no partials
        for (int c = 0; c < 6; c++) st->kin[c] = K[c] * ks2;
    if (st->want_obs) {
        st->pe = pe;
        for (int c = 0; c < 6; c++) st->vir[c] = W[c];
        st->obs_step = cur;
    }
    st->step = cur;
    st->rec_step = cur;
    const double Ld = (double)st->L, V = Ld * Ld * Ld;
    if (st->want_obs && st->thermo_freq > 0 && (cur % st->thermo_freq) == 0 && st->thermo_cap > 0) {
        ThermoRec r;
        const double N = (double)d.natoms;
        r.step = cur;
        r.ke = 0.5 * (st->kin[0] + st->kin[3] + st->kin[5]) / N;
        r.pe = pe / N;
        r.temp = 2.0 * r.ke / 3.0;  // src/analysis/thermo.cu:178
        for (int c = 0; c < 6; c++) r.press[c] = (st->kin[c] + W[c]) / V;
        r.L = Ld;
        float m2;
        int g;
        max_key_split(st->vkey, &m2, &g);
        r.maxv = sqrt((double)m2) * (double)st->v_pending;
        r.maxv_id = g;
        max_key_split(st->fkey, &m2, &g);
        r.maxf = sqrt((double)m2);
        r.maxf_id = g;
        d.thermo[st->thermo_written % st->thermo_cap] = r;
        st->thermo_written++;
    }
    // stress sample: the pressure tensor of the thermo record (src/main.cu:633,658-664)
    const int sample = d.corr && st->stress_freq > 0 && (cur % st->stress_freq) == 0;
    if (sample)
        for (int c = 0; c < 6; c++) smp[c] = (st->kin[c] + W[c]) / V;
    // the force pass of this step kicked ahead exactly when another step follows
    st->kicked = (d.ensemble == 0 && cur < st->stop) ? 1 : 0;
    // a neighbor overflow (list build) or an overstretched bond (force pass)
    // found during this step stops the replica at the end of it
    if (cur >= st->stop || st->err) {
        st->active = 0;
        st->pre_ready = 0;
    } else {
        start_step(d, st, cur + 1);
    }
    return sample;
}

// ---------------------------------------------------------------------------
// Per-replica boundary: fixed-order FP64 reduction of the wave partials, then
// the scalar integrator work and records (boundary_tail), then the stress
// correlator push on sample steps.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBndThreads) k_boundary(const RepDesc* __restrict__ descs,
                                                          const int* __restrict__ reps, int nr, int mode) {
    __shared__ double red[kNumSum + 2][kBndThreads / 64];
    __shared__ double smp[6];
    __shared__ int push;
    const RepDesc& d = descs[reps[blockIdx.x]];
    RepState* st = d.st;
    if (mode == BND_START) {
        if (threadIdx.x == 0 && st->active && !st->pre_ready) start_step(d, st, st->step + 1);
        return;
    }
    if (mode == BND_STEP && !st->active) return;
    // fixed-order reduction: per-thread strided sums, wave shuffles, then the
    // waves in order (the same order whatever else is in the batch)
    // second level: lane l of a wave reads component l % 16 of chunk 4 w + l / 16
    // (+ 4 x waves per pass), the chunks in a fixed order
    {
        // a Langevin step that records nothing reduced no partials (k_force)
        const bool none = mode == BND_STEP && d.ensemble == 0 && !st->want_obs;
        const int lane = lane_id(), wv = threadIdx.x >> 6, c = lane & 15;
        const bool mx = c >= kNumSum;  // the two maximum keys
        double a = 0.0;
        for (int ck = 4 * wv + (lane >> 4); !none && ck < d.nchunks; ck += 4 * (kBndThreads / 64)) {
            const double x = d.cpart[(size_t)ck * kNumPart + c];
            a = mx ? fmax(a, x) : a + x;
        }
        double o = __shfl_xor(a, 16, 64);
        a = mx ? fmax(a, o) : a + o;
        o = __shfl_xor(a, 32, 64);
        a = mx ? fmax(a, o) : a + o;
        if (lane < kNumSum + 2) red[lane][wv] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) push = boundary_tail(d, st, mode, red, smp);
    __syncthreads();
    if (push) corr_push(d.corr, smp);
}

// Reference cell id and Morton key by global id (parity entry points).
__global__ void k_ref_cells(const RepDesc* __restrict__ dp, int natoms, int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natoms) return;
    const RepDesc& d = *dp;
    const RepState* st = d.st;
    int n = (int)(st->Lp / d.rcs);
    const int nx = n < 1 ? 1 : n;
    const float ics = 1.0f / (st->Lp / (float)nx);
    const float4 r = d.pos[i];
    out[d.img[i].w] = (cell_coord(r.z, ics, nx) * nx + cell_coord(r.y, ics, nx)) * nx +
                      cell_coord(r.x, ics, nx);
}

#pragma clang fp contract(off)
__device__ inline uint32_t quant1024(float x, float invL) {
    const float q = x * invL * 1024.0f;
    const uint32_t u = (q > 0.0f) ? (uint32_t)q : 0u;  // truncation, saturating at 0
    return u > 1023u ? 1023u : u;
}
#pragma clang fp contract(fast)

__global__ void k_morton_keys(const RepDesc* __restrict__ dp, int natoms, uint32_t* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natoms) return;
    const RepDesc& d = *dp;
    const float invL = d.st->invLp;
    const float4 r = d.pos[i];
    out[d.img[i].w] = (expand_bits(quant1024(r.z, invL)) << 2) |
                      (expand_bits(quant1024(r.y, invL)) << 1) | expand_bits(quant1024(r.x, invL));
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int cdiv(int a, int b) { return (a + b - 1) / b; }
// grid cap of the grid-stride rebuild kernels: a step's rebuild list (about
// 4,000 tiles) in one pass, without thousands of waves that find no tile
static constexpr int kStrideBlocks = 1024;
static constexpr int kBuildBlocks = 8192;   // grid cap of the list-build workgroups (one group each)
static constexpr int kForceListBlocks = 1280;  // workgroups of a listed force pass (five per CU; a multiple of kXcds)
static_assert(kForceListBlocks % kXcds == 0, "the listed force pass splits its grid evenly over the XCDs");

void launch_pre(const Launch& l) {
    if (l.ntl <= 0) return;
    k_pre<<<cdiv(l.ntl, 4), 256, 0, l.stream>>>(l.descs, l.tiles, l.ntl, l.nrt);
}
void launch_decide(const Launch& l) {
    if (l.nr <= 0) return;
    k_decide<<<l.nr, 1024, 0, l.stream>>>(l.descs, l.reps, l.nr, l.rep_t0, l.rlist, l.nrt);
}
void launch_rg(const RepDesc* d, int nchains, double* out, hipStream_t s) {
    if (nchains <= 0) return;
    k_rg<<<nchains, 64, 0, s>>>(d, out);
}
void launch_rg_mean(const RepDesc* d, const double* vals, double* out, hipStream_t s) {
    k_rg_mean<<<1, 64, 0, s>>>(d, vals, out);
}
void launch_corr_push(const RepDesc* d, const Sample6& x, hipStream_t s) {
    k_corr_push<<<1, 128, 0, s>>>(d, x);
}
void launch_nh_chain_run(RepState* st, const double* ke, int n, int carry, double kT, int natoms,
                         double half_dt, double* scales, hipStream_t s) {
    if (n <= 0) return;
    k_nh_chain_run<<<1, 64, 0, s>>>(st, ke, n, carry, kT, natoms, half_dt, scales);
}
void launch_status(const Launch& l, int4* out) {
    if (l.nr <= 0) return;
    k_status<<<cdiv(l.nr, 256), 256, 0, l.stream>>>(l.descs, l.reps, l.nr, out);
}
void launch_phase(const Launch& l, int* out) {
    if (l.nr <= 0) return;
    k_phase<<<cdiv(l.nr, 256), 256, 0, l.stream>>>(l.descs, l.reps, l.nr, out);
}
void launch_control(const Launch& l, int op, int64_t arg) {
    if (l.nr <= 0) return;
    k_control<<<cdiv(l.nr, 64), 64, 0, l.stream>>>(l.descs, l.reps, l.nr, op, arg);
}
// k_build and k_force are compiled for at most kMaxBlock threads per block;
// any other request runs at the default shape.
static constexpr int kMaxBlock = kMaxBlockK;
static inline int valid_block(int b) { return (b == 64 || b == 128 || b == 256) ? b : kMaxBlock; }

void launch_rebuild(const Launch& l, int always, int build) {
    if (l.ntl <= 0) return;
    // a step's rebuild tile list is k_decide's; otherwise list the tiles here
    if (always) k_rebuild_list<<<1, 1024, 0, l.stream>>>(l.descs, l.reps, l.rep_t0, l.nr, always, l.rlist, l.nrt);
    int g = cdiv(l.ntl, 4);
    if (g > kStrideBlocks) g = kStrideBlocks;
    k_bin<<<g, 256, 0, l.stream>>>(l.descs, l.tiles, l.rlist, l.nrt);
    k_scan<<<l.nr, 1024, 0, l.stream>>>(l.descs, l.reps, l.nr, always);
    k_scatter<<<g, 256, 0, l.stream>>>(l.descs, l.tiles, l.rlist, l.nrt);
    k_rank<<<g, 256, 0, l.stream>>>(l.descs, l.tiles, l.rlist, l.nrt);
    k_gather<<<g, 256, 0, l.stream>>>(l.descs, l.tiles, l.rlist, l.nrt);
    if (build) launch_build(l);
}
void launch_build(const Launch& l, int max_blocks) {
    if (l.ntl <= 0) return;
    int gb = cdiv(l.ntl, kGroupTiles);
    const int cap = max_blocks > 0 ? max_blocks : kBuildBlocks;
    if (gb > cap) gb = cap;
    k_build<<<gb, kGroupSlots, 0, l.stream>>>(l.descs, l.tiles, l.rlist, l.nrt);
}
static void force_kernel(const Launch& l, int g, int block, int t0, int nt, int terms, int kick, int mode,
                         int sel) {
    const size_t lds = 0;
#define FMD_FORCE_ARGS l.descs, l.tiles, t0, nt, kick, mode, sel, l.rlist, l.nrt
    switch (terms & 7) {
        case 1: k_force<1><<<g, block, lds, l.stream>>>(FMD_FORCE_ARGS); break;
        case 2: k_force<2><<<g, block, lds, l.stream>>>(FMD_FORCE_ARGS); break;
        case 3: k_force<3><<<g, block, lds, l.stream>>>(FMD_FORCE_ARGS); break;
        case 4: k_force<4><<<g, block, lds, l.stream>>>(FMD_FORCE_ARGS); break;
        case 5: k_force<5><<<g, block, lds, l.stream>>>(FMD_FORCE_ARGS); break;
        case 6: k_force<6><<<g, block, lds, l.stream>>>(FMD_FORCE_ARGS); break;
        case 7: k_force<7><<<g, block, lds, l.stream>>>(FMD_FORCE_ARGS); break;
        default: k_force<0><<<g, block, lds, l.stream>>>(FMD_FORCE_ARGS); break;
    }
#undef FMD_FORCE_ARGS
}
void launch_force(const Launch& l, int t0, int nt, int block, int terms, int kick, int mode, int sel) {
    if (nt <= 0) return;
    block = valid_block(block);
    force_kernel(l, cdiv(nt, block / 64), block, t0, nt, terms, kick, mode, sel);
}
void launch_force_listed(const Launch& l, int kick, int mode) {
    if (l.ntl <= 0) return;
    k_force<7, true><<<kForceListBlocks, kGroupSlots, 0, l.stream>>>(l.descs, l.tiles, 0, 0, kick, mode, FS_ALL,
                                                                      l.rlist, l.nrt);
}
void launch_boundary(const Launch& l, int mode) {
    if (l.nr <= 0) return;
    if (mode != BND_START && l.nch > 0) k_chunk<<<l.nch, 256, 0, l.stream>>>(l.descs, l.chunks, l.nch, mode);
    k_boundary<<<l.nr, kBndThreads, 0, l.stream>>>(l.descs, l.reps, l.nr, mode);
}
void launch_ref_cells(const RepDesc* d, int natoms, int* out, hipStream_t s) {
    if (natoms <= 0) return;
    k_ref_cells<<<cdiv(natoms, 256), 256, 0, s>>>(d, natoms, out);
}
void launch_morton_keys(const RepDesc* d, int natoms, uint32_t* out, hipStream_t s) {
    if (natoms <= 0) return;
    k_morton_keys<<<cdiv(natoms, 256), 256, 0, s>>>(d, natoms, out);
}

}  // namespace fmd
