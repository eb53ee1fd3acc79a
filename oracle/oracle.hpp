// oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of fastMD's per-timestep Kremer-Grest algorithm (core, neighbor,
// force, integrate).  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this code, and only as the checker: the product
// library (fastmd_amd/libfastmd.so) never links it and has no CPU fallback.
//
// Two flavours of each force term:
//   * *_literal_f32 : a line-by-line restatement of reference/cpu_reference.hpp
//                     (FP32 throughout, same operation order), pinned against the
//                     reference's own unit-test inputs and closed forms.
//   * forces_f64    : the FP64 yardstick the GPU's FP32 force arithmetic is
//                     measured against.  Cutoff decisions are taken on the same
//                     FP32 squared distance the GPU computes (so both sides sum
//                     the same pair set); the arithmetic on each pair is FP64.
//
// Every function cites the reference file:line it follows.  Each is pinned by
// tests/test_oracle_pins.py against the reference's own test inputs, closed
// forms, brute force or an independent formulation.  The Langevin noise is a
// counter-based Philox4x32-10 pinned by its published known-answer vectors; the
// reference's cuRAND stream is NVIDIA library code and is not reproduced, so
// bit-level parity of the noise with the reference is unpinned (it is matched
// statistically).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Type id bit-cast into pos.w  (src/core/types.cuh:23-41)
// ---------------------------------------------------------------------------
inline int type_of(float w) { int t; std::memcpy(&t, &w, 4); return t; }
inline float type_bits(int t) { float w; std::memcpy(&w, &t, 4); return w; }

// ---------------------------------------------------------------------------
// PBC  (reference/cpu_reference.hpp:7-22, src/core/pbc.cuh:4-28)
// All float code in this file is compiled with -ffp-contract=off and without
// FMA, i.e. IEEE single precision with one rounding per operation.
// ---------------------------------------------------------------------------
inline float min_image_ref(float dr, float L, float invL) {
    return dr - L * std::rint(dr * invL);
}
inline float dist2_ref(const float* a, const float* b, float L, float invL) {
    float dx = min_image_ref(a[0] - b[0], L, invL);
    float dy = min_image_ref(a[1] - b[1], L, invL);
    float dz = min_image_ref(a[2] - b[2], L, invL);
    return dx * dx + dy * dy + dz * dz;
}
inline float wrap_ref(float x, float L, float invL) { return x - L * std::floor(x * invL); }

// The build's force-path displacement: each coordinate as a 32-bit fixed-point
// number X = floor(x * 2^32 / L) mod 2^32 (FP64 product), the two's-complement
// difference of two of them converted to FP32 and scaled by L.  Restated
// bit-for-bit so the oracle takes the same LJ cutoff decisions as the device.
inline uint32_t fix_coord(float x, double scale) {
    return (uint32_t)(long long)std::floor((double)x * scale);
}
inline float fix_diff(uint32_t a, uint32_t b, float L) {
    return (float)(int32_t)(a - b) * (L * 2.3283064365386963e-10f);  // L * 2^-32 is exact
}
inline float min_image_fixed_f(float a, float b, float L) {
    const double sc = 4294967296.0 / (double)L;
    return fix_diff(fix_coord(a, sc), fix_coord(b, sc), L);
}
// The device's LJ cutoff test in grid units (grid step u = L * 2^-32):
// a = (float)(int32)(Xi - Xj) per axis, s = fma(az, az, fma(ay, ay,
// fma(ax, ax, -lo_g))), inside when 0 < s < span, lo_g = (float)(1e-10f / u^2),
// span = (float)((rc2 - 1e-10f) / u^2), quotients in FP64.  Mathematically
// s u^2 = r^2 - 1e-10, so this is the r^2 < rc^2 && r^2 > 1e-10 test of
// src/force/lj.cu:37 up to one FP32 rounding of each side.
inline bool lj_inside_grid(const float* pi, const float* pj, float rc2, float L) {
    const double sc = 4294967296.0 / (double)L;
    const float u = L * 2.3283064365386963e-10f;
    const double u2 = (double)u * (double)u;
    const float lo_g = (float)((double)1e-10f / u2);
    const float span = (float)(((double)rc2 - (double)1e-10f) / u2);
    float a[3];
    for (int k = 0; k < 3; k++) a[k] = (float)(int32_t)(fix_coord(pi[k], sc) - fix_coord(pj[k], sc));
    const float s = std::fma(a[2], a[2], std::fma(a[1], a[1], std::fma(a[0], a[0], -lo_g)));
    return s > 0.0f && s < span;
}
// Same displacement carried in FP64 (the image shift m chosen as above).
inline double min_image_exact_d(float a, float b, float L, float invL) {
    float s = a - b;
    float m = std::rint(s * invL);
    return ((double)a - (double)b) - (double)L * (double)m;
}

// ---------------------------------------------------------------------------
// Morton keys  (reference/cpu_reference.hpp:24-53; src/core/morton.cu:12-28)
// ---------------------------------------------------------------------------
inline uint32_t expand_bits(uint32_t v) {
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
inline uint32_t quantize1024(float x, float invL) {
    float q = x * invL * 1024.0f;              // (x*invL)*1024, left to right
    uint32_t u = (q > 0.0f) ? (uint32_t)q : 0u;  // __float2uint_rz saturates at 0
    return u > 1023u ? 1023u : u;
}
inline uint32_t morton3d(float x, float y, float z, float invL) {
    return (expand_bits(quantize1024(z, invL)) << 2) |
           (expand_bits(quantize1024(y, invL)) << 1) | expand_bits(quantize1024(x, invL));
}
// Stable permutation by (key, original index) -- std::sort of pairs
// (reference/cpu_reference.hpp:42-53).
inline std::vector<int> morton_perm(const float* pos4, int n, float invL) {
    std::vector<std::pair<uint32_t, int>> k(n);
    for (int i = 0; i < n; i++)
        k[i] = {morton3d(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], invL), i};
    std::sort(k.begin(), k.end());
    std::vector<int> p(n);
    for (int i = 0; i < n; i++) p[i] = k[i].second;
    return p;
}

// ---------------------------------------------------------------------------
// Cell grid and cell index  (src/neighbor/verlet_list.cu:3-13, :58-75)
// ---------------------------------------------------------------------------
struct CellGrid {
    int nx;
    float cell_size, inv_cs;
};
inline CellGrid cell_grid(float L, float rcs) {
    CellGrid g;
    int n = (int)(L / rcs);
    g.nx = n < 1 ? 1 : n;
    g.cell_size = L / (float)g.nx;
    g.inv_cs = 1.0f / g.cell_size;
    return g;
}
inline void cell_coords(const float* p, const CellGrid& g, int* c) {
    for (int a = 0; a < 3; a++) {
        int ca = (int)std::floor(p[a] * g.inv_cs);   // __float2int_rd
        c[a] = ((ca % g.nx) + g.nx) % g.nx;
    }
}
inline int cell_id(const float* p, const CellGrid& g) {
    int c[3];
    cell_coords(p, g, c);
    return (c[2] * g.nx + c[1]) * g.nx + c[0];
}

// ---------------------------------------------------------------------------
// 1-2 exclusions from bonds, CSR, sorted unique  (src/io/lammps_data.cpp:231-253)
// ---------------------------------------------------------------------------
inline void build_exclusions(int n, const int* bonds, int nbonds, std::vector<int>& off,
                             std::vector<int>& lst) {
    std::vector<std::vector<int>> per(n);
    for (int b = 0; b < nbonds; b++) {
        per[bonds[2 * b]].push_back(bonds[2 * b + 1]);
        per[bonds[2 * b + 1]].push_back(bonds[2 * b]);
    }
    off.assign(n + 1, 0);
    lst.clear();
    for (int i = 0; i < n; i++) {
        auto& l = per[i];
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        off[i] = (int)lst.size();
        lst.insert(lst.end(), l.begin(), l.end());
    }
    off[n] = (int)lst.size();
}
inline bool excluded(const int* off, const int* lst, int i, int j) {
    if (!off) return false;
    for (int e = off[i]; e < off[i + 1]; e++)
        if (lst[e] == j) return true;
    return false;
}

// ---------------------------------------------------------------------------
// Verlet pair set: ordered pairs (i,j), i!=j, dist2_ref < rcs^2, not excluded,
// no cap.  Brute force over all pairs, the criterion of
// src/neighbor/verlet_list.cu:259-286 with reference/cpu_reference.hpp:11-18.
// ---------------------------------------------------------------------------
inline void pair_set_brute(const float* pos4, int n, float L, float invL, float rcs,
                           const int* eoff, const int* elst, std::vector<int>& out) {
    float rcs2 = rcs * rcs;
    out.clear();
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            if (i == j) continue;
            if (dist2_ref(pos4 + 4 * i, pos4 + 4 * j, L, invL) < rcs2 &&
                !excluded(eoff, elst, i, j)) {
                out.push_back(i);
                out.push_back(j);
            }
        }
}

// Same set, enumerated through the cell grid with the 27-cell stencil and the
// visited-cell de-duplication of src/neighbor/verlet_list.cu:154-223.  Output is
// sorted by (i, j).  Used where brute force is too slow; checked equal to
// pair_set_brute on small systems.
inline void pair_set_cells(const float* pos4, int n, float L, float invL, float rcs,
                           const int* eoff, const int* elst, std::vector<int>& out) {
    CellGrid g = cell_grid(L, rcs);
    int nc = g.nx * g.nx * g.nx;
    std::vector<int> head(nc + 1, 0), cid(n), items(n);
    for (int i = 0; i < n; i++) { cid[i] = cell_id(pos4 + 4 * i, g); head[cid[i] + 1]++; }
    for (int c = 0; c < nc; c++) head[c + 1] += head[c];
    std::vector<int> fill(head.begin(), head.end() - 1);
    for (int i = 0; i < n; i++) items[fill[cid[i]]++] = i;
    float rcs2 = rcs * rcs;
    out.clear();
    std::vector<int> nb;
    for (int i = 0; i < n; i++) {
        int c[3];
        cell_coords(pos4 + 4 * i, g, c);
        int visited[27], nv = 0;
        nb.clear();
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int x = (c[0] + dx + g.nx) % g.nx, y = (c[1] + dy + g.nx) % g.nx,
                        z = (c[2] + dz + g.nx) % g.nx;
                    int cc = (z * g.nx + y) * g.nx + x;
                    bool seen = false;
                    for (int v = 0; v < nv; v++) seen |= (visited[v] == cc);
                    if (seen) continue;
                    visited[nv++] = cc;
                    for (int t = head[cc]; t < head[cc + 1]; t++) {
                        int j = items[t];
                        if (j == i) continue;
                        if (dist2_ref(pos4 + 4 * i, pos4 + 4 * j, L, invL) < rcs2 &&
                            !excluded(eoff, elst, i, j))
                            nb.push_back(j);
                    }
                }
        std::sort(nb.begin(), nb.end());
        for (int j : nb) { out.push_back(i); out.push_back(j); }
    }
}

// ---------------------------------------------------------------------------
// Literal FP32 restatements of reference/cpu_reference.hpp force terms.
// Outputs: f4[4*i] = (fx, fy, fz, pe_i); vir[6] = xx, xy, xz, yy, yz, zz.
// ---------------------------------------------------------------------------
// reference/cpu_reference.hpp:95-153 (brute-force LJ, FP32 accumulators)
inline void lj_literal_f32(const float* pos4, int n, const float* lj2, int ntypes, float rc,
                           float L, float invL, const int* eoff, const int* elst, float* f4,
                           float* vir) {
    float rc2 = rc * rc;
    for (int i = 0; i < 4 * n; i++) f4[i] = 0.0f;
    for (int k = 0; k < 6; k++) vir[k] = 0.0f;
    for (int i = 0; i < n; i++) {
        int ti = type_of(pos4[4 * i + 3]);
        for (int j = 0; j < n; j++) {
            if (i == j) continue;
            if (excluded(eoff, elst, i, j)) continue;
            int tj = type_of(pos4[4 * j + 3]);
            float dx = min_image_ref(pos4[4 * i] - pos4[4 * j], L, invL);
            float dy = min_image_ref(pos4[4 * i + 1] - pos4[4 * j + 1], L, invL);
            float dz = min_image_ref(pos4[4 * i + 2] - pos4[4 * j + 2], L, invL);
            float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= rc2) continue;
            float eps = lj2[2 * (ti * ntypes + tj)], sig = lj2[2 * (ti * ntypes + tj) + 1];
            float sig2 = sig * sig;
            float r2inv = 1.0f / r2;
            float sr2 = sig2 * r2inv;
            float sr6 = sr2 * sr2 * sr2;
            float sr12 = sr6 * sr6;
            float force_r = 24.0f * eps * r2inv * (2.0f * sr12 - sr6);
            float pair_pe = 4.0f * eps * (sr12 - sr6);
            f4[4 * i] += force_r * dx;
            f4[4 * i + 1] += force_r * dy;
            f4[4 * i + 2] += force_r * dz;
            f4[4 * i + 3] += 0.5f * pair_pe;
            vir[0] += 0.5f * force_r * dx * dx;
            vir[1] += 0.5f * force_r * dx * dy;
            vir[2] += 0.5f * force_r * dx * dz;
            vir[3] += 0.5f * force_r * dy * dy;
            vir[4] += 0.5f * force_r * dy * dz;
            vir[5] += 0.5f * force_r * dz * dz;
        }
    }
}

// reference/cpu_reference.hpp:161-225 (FENE + WCA per bond; bp = k,R0,eps,sig)
inline void fene_literal_f32(const float* pos4, int n, const int* bonds, const int* btype,
                             int nbonds, const float* bp, float L, float invL, float* f4,
                             float* vir) {
    for (int i = 0; i < 4 * n; i++) f4[i] = 0.0f;
    for (int k = 0; k < 6; k++) vir[k] = 0.0f;
    for (int b = 0; b < nbonds; b++) {
        int i = bonds[2 * b], j = bonds[2 * b + 1], t = btype[b];
        float k = bp[4 * t], R0 = bp[4 * t + 1], eps = bp[4 * t + 2], sig = bp[4 * t + 3];
        float dx = min_image_ref(pos4[4 * i] - pos4[4 * j], L, invL);
        float dy = min_image_ref(pos4[4 * i + 1] - pos4[4 * j + 1], L, invL);
        float dz = min_image_ref(pos4[4 * i + 2] - pos4[4 * j + 2], L, invL);
        float r2 = dx * dx + dy * dy + dz * dz;
        float R02 = R0 * R0;
        float fene_ff = -k * R02 / (R02 - r2);
        float sig2 = sig * sig;
        float r_cut2 = sig2 * 1.2599210498948732f;
        float wca_ff = 0.0f, wca_pe = 0.0f;
        if (r2 < r_cut2) {
            float r2inv = 1.0f / r2;
            float sr2 = sig2 * r2inv;
            float sr6 = sr2 * sr2 * sr2;
            float sr12 = sr6 * sr6;
            wca_ff = 24.0f * eps * r2inv * (2.0f * sr12 - sr6);
            wca_pe = 4.0f * eps * (sr12 - sr6) + eps;
        }
        float total_ff = fene_ff + wca_ff;
        float fix = total_ff * dx, fiy = total_ff * dy, fiz = total_ff * dz;
        f4[4 * i] += fix; f4[4 * i + 1] += fiy; f4[4 * i + 2] += fiz;
        f4[4 * j] -= fix; f4[4 * j + 1] -= fiy; f4[4 * j + 2] -= fiz;
        float fene_pe = -0.5f * k * R02 * std::log(1.0f - r2 / R02);
        float pair_pe = fene_pe + wca_pe;
        f4[4 * i + 3] += 0.5f * pair_pe;
        f4[4 * j + 3] += 0.5f * pair_pe;
        vir[0] += fix * dx; vir[1] += fix * dy; vir[2] += fix * dz;
        vir[3] += fiy * dy; vir[4] += fiy * dz; vir[5] += fiz * dz;
    }
}

// reference/cpu_reference.hpp:233-298 (harmonic angle in theta; ap = k, theta0[rad])
inline void angle_literal_f32(const float* pos4, int n, const int* ang4, int nang,
                              const float* ap, float L, float invL, float* f4, float* vir) {
    for (int i = 0; i < 4 * n; i++) f4[i] = 0.0f;
    for (int k = 0; k < 6; k++) vir[k] = 0.0f;
    for (int a = 0; a < nang; a++) {
        int i = ang4[4 * a], j = ang4[4 * a + 1], k = ang4[4 * a + 2], t = ang4[4 * a + 3];
        float ka = ap[2 * t], theta0 = ap[2 * t + 1];
        float dxij = min_image_ref(pos4[4 * i] - pos4[4 * j], L, invL);
        float dyij = min_image_ref(pos4[4 * i + 1] - pos4[4 * j + 1], L, invL);
        float dzij = min_image_ref(pos4[4 * i + 2] - pos4[4 * j + 2], L, invL);
        float dxkj = min_image_ref(pos4[4 * k] - pos4[4 * j], L, invL);
        float dykj = min_image_ref(pos4[4 * k + 1] - pos4[4 * j + 1], L, invL);
        float dzkj = min_image_ref(pos4[4 * k + 2] - pos4[4 * j + 2], L, invL);
        float rij2 = dxij * dxij + dyij * dyij + dzij * dzij;
        float rkj2 = dxkj * dxkj + dykj * dykj + dzkj * dzkj;
        float rij = std::sqrt(rij2), rkj = std::sqrt(rkj2);
        float c = (dxij * dxkj + dyij * dykj + dzij * dzkj) / (rij * rkj);
        c = std::fmin(std::fmax(c, -1.0f), 1.0f);
        float theta = std::acos(c);
        float dtheta = theta - theta0;
        float s = std::sqrt(1.0f - c * c);
        if (s < 1e-6f) s = 1e-6f;
        float pref = 2.0f * ka * dtheta / s;
        float fix = pref * (dxkj / (rij * rkj) - c * dxij / rij2);
        float fiy = pref * (dykj / (rij * rkj) - c * dyij / rij2);
        float fiz = pref * (dzkj / (rij * rkj) - c * dzij / rij2);
        float fkx = pref * (dxij / (rij * rkj) - c * dxkj / rkj2);
        float fky = pref * (dyij / (rij * rkj) - c * dykj / rkj2);
        float fkz = pref * (dzij / (rij * rkj) - c * dzkj / rkj2);
        f4[4 * i] += fix; f4[4 * i + 1] += fiy; f4[4 * i + 2] += fiz;
        f4[4 * k] += fkx; f4[4 * k + 1] += fky; f4[4 * k + 2] += fkz;
        f4[4 * j] -= (fix + fkx); f4[4 * j + 1] -= (fiy + fky); f4[4 * j + 2] -= (fiz + fkz);
        float pe = ka * dtheta * dtheta;
        f4[4 * i + 3] += pe / 3.0f; f4[4 * j + 3] += pe / 3.0f; f4[4 * k + 3] += pe / 3.0f;
        vir[0] += fix * dxij + fkx * dxkj;
        vir[1] += fix * dyij + fkx * dykj;
        vir[2] += fix * dzij + fkx * dzkj;
        vir[3] += fiy * dyij + fky * dykj;
        vir[4] += fiy * dzij + fky * dzkj;
        vir[5] += fiz * dzij + fkz * dzkj;
    }
}

// ---------------------------------------------------------------------------
// FP64 yardstick for the device force kernel.
// terms: bit 0 = LJ (src/force/lj.cu:27-63) and tabulated pairs, bit 1 = FENE+WCA
// (src/force/fene.cu:21-65), bit 2 = angle (src/force/angle.cu:22-79).
// LJ pairs: i != j, not excluded, and inside by the device's grid-unit test
// (lj_inside_grid); the force/r clamp of lj.cu:48 is
// applied when clamp_lj is set.  FENE r^2 is clamped to [1e-6, 0.9999 R0^2]
// as in fene.cu:33 when clamp_fene is set.  Angles use theta = atan2(|a x b|, a.b),
// mathematically identical to acos(clamp(cos)) of angle.cu:37-45.
// fcap > 0 caps the LJ pair force magnitude |f| <= fcap (push-off only).
// f8 per atom (double): fx, fy, fz, pe; vir[6] double.
// ---------------------------------------------------------------------------
struct ForceParams {
    int ntypes;
    const float* lj2;  // ntypes^2 x (eps, sig)
    float rc;
    const float* bp;   // per bond type: k, R0, eps, sig
    const float* ap;   // per angle type: k, theta0
    int clamp_lj, clamp_fene;
    float fcap;
    // tabulated pairs: tab_idx[ntypes^2] (-1 = none), per table its point
    // count and (rmin, rmax), rows (r, energy, force/r) of all tables in order
    int ntables = 0;
    const int* tab_idx = nullptr;
    const int* tab_n = nullptr;
    const float* tab_range = nullptr;
    const float* tab_data = nullptr;
};

// Linear interpolation of one table in FP64 (src/force/table.cu:87-104): the
// nodes sit at rmin + k dr with dr = (rmax - rmin) / (n - 1); r is clamped to
// [rmin, rmax], k = floor((r - rmin) / dr) clamped to [0, n - 2], and
// t = (r - rmin) / dr - k weights nodes k and k + 1.  rows hold (r, e, f).
inline void table_eval_f64(int n, float rmin, float rmax, const float* rows, double r, double* f,
                           double* e) {
    const double lo = rmin, hi = rmax, dr = (hi - lo) / (double)(n - 1);
    if (r < lo) r = lo;
    if (r > hi) r = hi;
    const double x = (r - lo) / dr;
    int k = (int)std::floor(x);
    if (k < 0) k = 0;
    if (k > n - 2) k = n - 2;
    const double t = x - (double)k;
    const float* a = rows + 3 * (size_t)k;
    const float* b = a + 3;
    *f = (double)a[2] + t * ((double)b[2] - (double)a[2]);
    *e = (double)a[1] + t * ((double)b[1] - (double)a[1]);
}

inline void lj_pair_f64(const float* pi, const float* pj, const ForceParams& P, float L,
                        float invL, double* fi, double* pe_i, double* vir) {
    float rc2 = P.rc * P.rc;
    if (!lj_inside_grid(pi, pj, rc2, L)) return;
    double dx = min_image_exact_d(pi[0], pj[0], L, invL);
    double dy = min_image_exact_d(pi[1], pj[1], L, invL);
    double dz = min_image_exact_d(pi[2], pj[2], L, invL);
    double r2 = dx * dx + dy * dy + dz * dz;
    int ti = type_of(pi[3]), tj = type_of(pj[3]);
    double eps = P.lj2[2 * (ti * P.ntypes + tj)], sig = P.lj2[2 * (ti * P.ntypes + tj) + 1];
    double sr2 = sig * sig / r2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
    double fr = 24.0 * eps / r2 * (2.0 * sr12 - sr6);
    if (P.clamp_lj) fr = std::fmin(std::fmax(fr, -1000.0), 1000.0);
    if (P.fcap > 0.0f) {
        double r = std::sqrt(r2), cap = (double)P.fcap / r;
        if (fr > cap) fr = cap;
        if (fr < -cap) fr = -cap;
    }
    fi[0] += fr * dx; fi[1] += fr * dy; fi[2] += fr * dz;
    *pe_i += 0.5 * 4.0 * eps * (sr12 - sr6);
    vir[0] += 0.5 * fr * dx * dx; vir[1] += 0.5 * fr * dx * dy; vir[2] += 0.5 * fr * dx * dz;
    vir[3] += 0.5 * fr * dy * dy; vir[4] += 0.5 * fr * dy * dz; vir[5] += 0.5 * fr * dz * dz;
}

// Tabulated pair (src/force/table.cu:58-118): selected like an LJ pair
// (lj_inside_grid); half the energy and half the virial go to atom i.
inline void table_pair_f64(const float* pi, const float* pj, const ForceParams& P,
                           const std::vector<size_t>& toff, float L, float invL, double* fi,
                           double* pe_i, double* vir) {
    int ti = type_of(pi[3]), tj = type_of(pj[3]);
    int tb = P.tab_idx[ti * P.ntypes + tj];
    if (tb < 0) return;
    if (!lj_inside_grid(pi, pj, P.rc * P.rc, L)) return;
    double dx = min_image_exact_d(pi[0], pj[0], L, invL);
    double dy = min_image_exact_d(pi[1], pj[1], L, invL);
    double dz = min_image_exact_d(pi[2], pj[2], L, invL);
    double fr, e;
    table_eval_f64(P.tab_n[tb], P.tab_range[2 * tb], P.tab_range[2 * tb + 1],
                   P.tab_data + 3 * toff[tb], std::sqrt(dx * dx + dy * dy + dz * dz), &fr, &e);
    fi[0] += fr * dx; fi[1] += fr * dy; fi[2] += fr * dz;
    *pe_i += 0.5 * e;
    vir[0] += 0.5 * fr * dx * dx; vir[1] += 0.5 * fr * dx * dy; vir[2] += 0.5 * fr * dx * dz;
    vir[3] += 0.5 * fr * dy * dy; vir[4] += 0.5 * fr * dy * dz; vir[5] += 0.5 * fr * dz * dz;
}

// Bond i-j: adds force on i and j, half the PE to each, full virial.
inline void fene_bond_f64(const float* pos4, int i, int j, int t, const ForceParams& P,
                          float L, float invL, double* f8, double* vir) {
    double k = P.bp[4 * t], R0 = P.bp[4 * t + 1], eps = P.bp[4 * t + 2], sig = P.bp[4 * t + 3];
    double dx = min_image_exact_d(pos4[4 * i], pos4[4 * j], L, invL);
    double dy = min_image_exact_d(pos4[4 * i + 1], pos4[4 * j + 1], L, invL);
    double dz = min_image_exact_d(pos4[4 * i + 2], pos4[4 * j + 2], L, invL);
    double r2 = dx * dx + dy * dy + dz * dz;
    double R02 = R0 * R0;
    if (P.clamp_fene) r2 = std::fmin(std::fmax(r2, 1e-6), R02 * 0.9999);
    double fene_ff = -k * R02 / (R02 - r2);
    double sig2 = sig * sig, wca_ff = 0.0, wca_pe = 0.0;
    if (r2 < sig2 * 1.2599210498948732) {
        double sr2 = sig2 / r2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
        wca_ff = 24.0 * eps / r2 * (2.0 * sr12 - sr6);
        wca_pe = 4.0 * eps * (sr12 - sr6) + eps;
    }
    double ff = fene_ff + wca_ff;
    double fx = ff * dx, fy = ff * dy, fz = ff * dz;
    f8[4 * i] += fx; f8[4 * i + 1] += fy; f8[4 * i + 2] += fz;
    f8[4 * j] -= fx; f8[4 * j + 1] -= fy; f8[4 * j + 2] -= fz;
    double pe = -0.5 * k * R02 * std::log(1.0 - r2 / R02) + wca_pe;
    f8[4 * i + 3] += 0.5 * pe; f8[4 * j + 3] += 0.5 * pe;
    vir[0] += fx * dx; vir[1] += fx * dy; vir[2] += fx * dz;
    vir[3] += fy * dy; vir[4] += fy * dz; vir[5] += fz * dz;
}

inline void angle_f64(const float* pos4, int i, int j, int k, int t, const ForceParams& P,
                      float L, float invL, double* f8, double* vir) {
    double ka = P.ap[2 * t], th0 = P.ap[2 * t + 1];
    double a[3], b[3];
    for (int d = 0; d < 3; d++) {
        a[d] = min_image_exact_d(pos4[4 * i + d], pos4[4 * j + d], L, invL);
        b[d] = min_image_exact_d(pos4[4 * k + d], pos4[4 * j + d], L, invL);
    }
    double ra2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    double rb2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    double ra = std::sqrt(ra2), rb = std::sqrt(rb2);
    double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2],
           cz = a[0] * b[1] - a[1] * b[0];
    double cr = std::sqrt(cx * cx + cy * cy + cz * cz);
    double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    double theta = std::atan2(cr, dot);
    double c = dot / (ra * rb), s = cr / (ra * rb);
    if (s < 1e-6) s = 1e-6;
    double dth = theta - th0;
    double pref = 2.0 * ka * dth / s;
    double fi[3], fk[3];
    for (int d = 0; d < 3; d++) {
        fi[d] = pref * (b[d] / (ra * rb) - c * a[d] / ra2);
        fk[d] = pref * (a[d] / (ra * rb) - c * b[d] / rb2);
    }
    for (int d = 0; d < 3; d++) {
        f8[4 * i + d] += fi[d];
        f8[4 * k + d] += fk[d];
        f8[4 * j + d] -= fi[d] + fk[d];
    }
    double pe = ka * dth * dth;
    f8[4 * i + 3] += pe / 3.0; f8[4 * j + 3] += pe / 3.0; f8[4 * k + 3] += pe / 3.0;
    vir[0] += fi[0] * a[0] + fk[0] * b[0];
    vir[1] += fi[0] * a[1] + fk[0] * b[1];
    vir[2] += fi[0] * a[2] + fk[0] * b[2];
    vir[3] += fi[1] * a[1] + fk[1] * b[1];
    vir[4] += fi[1] * a[2] + fk[1] * b[2];
    vir[5] += fi[2] * a[2] + fk[2] * b[2];
}

// Full force evaluation in FP64.  LJ pairs are enumerated from the cell grid at
// rc (pair_set_cells criterion with rcs = rc): a superset test on FP32 reference
// distance with a margin, then the exact-difference decision in lj_pair_f64.
inline void forces_f64(const float* pos4, int n, float L, const ForceParams& P, int terms,
                       const int* eoff, const int* elst, const int* bonds, const int* btype,
                       int nbonds, const int* ang4, int nang, double* f8, double* vir) {
    float invL = 1.0f / L;
    for (int i = 0; i < 4 * n; i++) f8[i] = 0.0;
    for (int k = 0; k < 6; k++) vir[k] = 0.0;
    if (terms & 1) {
        std::vector<int> pairs;
        pair_set_cells(pos4, n, L, invL, P.rc * 1.0001f + 1e-4f, eoff, elst, pairs);
        std::vector<size_t> toff(P.ntables + 1, 0);
        for (int k = 0; k < P.ntables; k++) toff[k + 1] = toff[k] + (size_t)P.tab_n[k];
        for (size_t p = 0; p < pairs.size(); p += 2) {
            int i = pairs[p], j = pairs[p + 1];
            if (type_of(pos4[4 * i + 3]) < 0 || type_of(pos4[4 * j + 3]) < 0) continue;
            lj_pair_f64(pos4 + 4 * i, pos4 + 4 * j, P, L, invL, f8 + 4 * i, f8 + 4 * i + 3, vir);
            if (P.ntables)
                table_pair_f64(pos4 + 4 * i, pos4 + 4 * j, P, toff, L, invL, f8 + 4 * i,
                               f8 + 4 * i + 3, vir);
        }
    }
    if (terms & 2)
        for (int b = 0; b < nbonds; b++)
            fene_bond_f64(pos4, bonds[2 * b], bonds[2 * b + 1], btype[b], P, L, invL, f8, vir);
    if (terms & 4)
        for (int a = 0; a < nang; a++)
            angle_f64(pos4, ang4[4 * a], ang4[4 * a + 1], ang4[4 * a + 2], ang4[4 * a + 3], P, L,
                      invL, f8, vir);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 counter-based generator, from Salmon, Moraes, Dror & Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11.  Replaces the cuRAND
// Philox state of src/integrate/langevin.cu:7-28,64-66 with a stateless draw
// keyed by (replica seed, global bead id, step).
// ---------------------------------------------------------------------------
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += W0; k1 += W1; }
        uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0; c1 = lo1; c2 = hi0 ^ c3 ^ k1; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
// Uniform in (0,1) from the top 24 bits: exact in FP32 on host and device.
inline float u01(uint32_t x) { return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f); }
// Three standard normals for bead gid at step (Box-Muller, textbook form).
inline void normal3(uint64_t seed, uint32_t gid, int64_t step, float z[3]) {
    uint32_t ctr[4] = {gid, (uint32_t)step, (uint32_t)((uint64_t)step >> 32), 0x4C414E47u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    const float twopi = 6.28318530717958647692f;
    float a = std::sqrt(-2.0f * std::log(u01(r[0])));
    float b = std::sqrt(-2.0f * std::log(u01(r[2])));
    z[0] = a * std::cos(twopi * u01(r[1]));
    z[1] = a * std::sin(twopi * u01(r[1]));
    z[2] = b * std::cos(twopi * u01(r[3]));
}

// ---------------------------------------------------------------------------
// Nose-Hoover chain, one application (src/integrate/nose_hoover.cu:356-447,
// host nh_propagate_chain; device copy :86-178).  Templated on the scalar type:
// Real=float with nh_expf is the literal reference; Real=double uses exp().
// ---------------------------------------------------------------------------
template <typename Real> struct NHChain {
    int M;
    Real Q1, Qr;
    Real xi[10], vxi[10];
};
inline float nh_expf(float x) {  // src/integrate/nose_hoover.cu:72-84
    float x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x, x6 = x5 * x;
    return 1.0f + x + x2 * 0.5f + x3 * (1.0f / 6.0f) + x4 * (1.0f / 24.0f) +
           x5 * (1.0f / 120.0f) + x6 * (1.0f / 720.0f);
}
inline float chain_exp(float x) { return nh_expf(x); }
inline double chain_exp(double x) { return std::exp(x); }

template <typename Real>
Real nh_chain_step(NHChain<Real>& c, Real KE, Real kT, int natoms, Real half_dt) {
    const int M = c.M;
    Real Nf = (Real)3 * (Real)natoms;
    Real dt2 = half_dt, dt4 = half_dt * (Real)0.5, dt8 = half_dt * (Real)0.25;
    Real dd[10] = {};
    Real T = ((Real)2 * KE) / Nf;
    if (c.Q1 > (Real)0) dd[0] = (Nf * T - Nf * kT) / c.Q1;
    for (int ich = M - 1; ich > 0; ich--) {
        Real e = (ich + 1 < M) ? chain_exp(-dt8 * c.vxi[ich + 1]) : (Real)1;
        c.vxi[ich] *= e;
        c.vxi[ich] += dd[ich] * dt4;
        c.vxi[ich] *= e;
    }
    Real e0 = (M > 1) ? chain_exp(-dt8 * c.vxi[1]) : (Real)1;
    c.vxi[0] *= e0;
    c.vxi[0] += dd[0] * dt4;
    c.vxi[0] *= e0;
    Real scale = chain_exp(-dt2 * c.vxi[0]);
    T *= scale * scale;
    if (c.Q1 > (Real)0) dd[0] = (Nf * T - Nf * kT) / c.Q1;
    for (int ich = 0; ich < M; ich++) c.xi[ich] += dt2 * c.vxi[ich];
    c.vxi[0] *= e0;
    c.vxi[0] += dd[0] * dt4;
    c.vxi[0] *= e0;
    for (int ich = 1; ich < M; ich++) {
        Real e = (ich + 1 < M) ? chain_exp(-dt8 * c.vxi[ich + 1]) : (Real)1;
        c.vxi[ich] *= e;
        Real Qp = (ich == 1) ? c.Q1 : c.Qr;
        dd[ich] = (Qp * c.vxi[ich - 1] * c.vxi[ich - 1] - kT) / c.Qr;
        c.vxi[ich] += dd[ich] * dt4;
        c.vxi[ich] *= e;
    }
    return scale;
}

// Chain energy of the reference's conserved quantity (tests/test_nose_hoover.cu:19-40):
// 1/2 Q1 v0^2 + Nf kT xi0 + sum_{k>=1} (1/2 Qr vk^2 + kT xik).
template <typename Real> Real nh_chain_energy(const NHChain<Real>& c, Real kT, int natoms) {
    Real Nf = (Real)3 * natoms;
    Real e = (Real)0.5 * c.Q1 * c.vxi[0] * c.vxi[0] + Nf * kT * c.xi[0];
    for (int k = 1; k < c.M; k++) e += (Real)0.5 * c.Qr * c.vxi[k] * c.vxi[k] + kT * c.xi[k];
    return e;
}

// Chain energy that the dynamics of nh_chain_step actually conserves.
// Within one application the backward sweep gives the elements k >= 1 no
// force term (their acceleration array starts at zero and is only filled in
// the forward sweep, src/integrate/nose_hoover.cu:365,385-393,431-441), so in
// the small-dt limit element k >= 1 obeys
//   Qr dvk/dt = 1/2 (Q_{k-1} v_{k-1}^2 - kT) - Qr vk v_{k+1}
// with half the force of a textbook chain, while element 0 obeys
//   Q1 dv0/dt = (2 KE - Nf kT) - Q1 v0 v1  and particles dv/dt = F/m - v0 v.
// d/dt of KE + U + 1/2 Q1 v0^2 + Nf kT xi0 + sum_k (a_k 1/2 Qr vk^2 + b_k xik)
// vanishes term by term iff a_1 = 2, a_k = 2 a_{k-1} and b_k = a_k kT / 2, i.e.
// weights a_k = 2^k, b_k = 2^(k-1) kT.  For M = 1 both forms are equal.
template <typename Real> Real nh_chain_energy_dyn(const NHChain<Real>& c, Real kT, int natoms) {
    Real Nf = (Real)3 * natoms;
    Real e = (Real)0.5 * c.Q1 * c.vxi[0] * c.vxi[0] + Nf * kT * c.xi[0];
    Real a = (Real)1;
    for (int k = 1; k < c.M; k++) {
        a *= (Real)2;
        e += a * ((Real)0.5 * c.Qr * c.vxi[k] * c.vxi[k]) + (Real)0.5 * a * kT * c.xi[k];
    }
    return e;
}

// ---------------------------------------------------------------------------
// NPT barostat + chain (src/integrate/nose_hoover.cu:185-305).  advance = the
// post-force call (use_carry=false): three Suzuki-Yoshida substeps on eps and
// v_eps, then V, L and the barostat velocity factors, then the chain on the
// barostat-scaled KE.  Returns nh scale; fills the derived quantities.
// ---------------------------------------------------------------------------
template <typename Real> struct Baro {
    Real eps, veps, W, V0, Ptarget;
};
template <typename Real> struct BaroOut {
    Real V, L, baro_scale, exp_vW, veps_W_dt, nh_scale, KE_scaled;
};
template <typename Real>
BaroOut<Real> npt_chain_baro(NHChain<Real>& c, Baro<Real>& b, Real KE, Real vir_trace,
                             bool advance, Real kT, int natoms, Real half_dt, Real dt) {
    BaroOut<Real> o;
    Real Nf = (Real)3 * (Real)natoms;
    Real V = b.V0 * std::exp((Real)3 * b.eps);
    if (advance) {
        Real P = ((Real)2 * KE + vir_trace) / ((Real)3 * V);
        Real cb = std::cbrt((Real)2);
        Real w[3] = {(Real)1 / ((Real)2 - cb), -cb / ((Real)2 - cb), (Real)1 / ((Real)2 - cb)};
        for (int sy = 0; sy < 3; sy++) {
            Real ws = w[sy] * dt;
            Real dv = (Real)3 * V * (P - b.Ptarget) * ws;
            Real vh = b.veps + (Real)0.5 * dv;
            b.eps += vh * ws / b.W;
            dv = (Real)3 * b.V0 * std::exp((Real)3 * b.eps) * (P - b.Ptarget) * ws;
            b.veps += dv;
        }
        V = b.V0 * std::exp((Real)3 * b.eps);
    }
    o.V = V;
    o.L = std::cbrt(V);
    Real vW = b.veps / b.W;
    o.baro_scale = std::exp(-((Real)1 + (Real)3 / Nf) * vW * half_dt);
    o.exp_vW = std::exp(vW * dt);
    o.veps_W_dt = vW * dt;
    Real KEs = KE * o.baro_scale * o.baro_scale;
    o.KE_scaled = KEs;
    o.nh_scale = nh_chain_step(c, KEs, kT, natoms, half_dt);
    return o;
}
// sinh(x)/x with the small-argument branch of src/integrate/nose_hoover.cu:540-543
template <typename Real> Real sinch(Real x) {
    if (std::fabs(x) < (Real)1e-6) return (Real)1 + x * x / (Real)6;
    return std::sinh(x) / x;
}


// ---------------------------------------------------------------------------
// Radius of gyration (src/analysis/rg.cu:3-69, :106-129): chains are runs of
// equal molecule id in atom order; each chain's Rg is sqrt(mean |x - com|^2)
// over unwrapped positions x + image * L; the result is the chain average.
// FP64 throughout.
// ---------------------------------------------------------------------------
inline double rg_ensemble(int n, const float* pos4, const int* image, const int* mol, float L) {
    double sum = 0.0;
    int nch = 0;
    for (int s = 0; s < n;) {
        int e = s + 1;
        while (e < n && mol[e] == mol[s]) e++;
        double c[3] = {0.0, 0.0, 0.0};
        for (int i = s; i < e; i++)
            for (int k = 0; k < 3; k++) c[k] += (double)pos4[4 * i + k] + (double)image[3 * i + k] * (double)L;
        for (int k = 0; k < 3; k++) c[k] /= (double)(e - s);
        double r2 = 0.0;
        for (int i = s; i < e; i++)
            for (int k = 0; k < 3; k++) {
                const double x = (double)pos4[4 * i + k] + (double)image[3 * i + k] * (double)L - c[k];
                r2 += x * x;
            }
        sum += std::sqrt(r2 / (double)(e - s));
        nch++;
        s = e;
    }
    return nch ? sum / nch : 0.0;
}

// ---------------------------------------------------------------------------
// Multiple-tau correlator (src/analysis/correlator.cu:33-92), FP64 sums.
// Level l keeps a ring of P samples; a push writes the sample at the ring's
// insert position idx and adds sample * (the sample lag pushes back) to
// corr[l][lag] for lag = 0..idx (the reference's min(idx + 1, P) stored count,
// which restarts when the ring wraps), bumping counts[l][lag].  Every P pushes
// into a level, the mean of those P samples is pushed into the next level.
// ---------------------------------------------------------------------------
struct MultiTau {
    static constexpr int LV = 16, P = 16;
    std::vector<double> buf, corr, acc;
    std::vector<int> counts, ins, bcount;
    MultiTau() : buf(LV * P * 6, 0.0), corr(LV * P * 6, 0.0), acc(LV * 6, 0.0), counts(LV * P, 0),
                 ins(LV, 0), bcount(LV, 0) {}
    void push(const double* x_in) {
        double x[6];
        for (int c = 0; c < 6; c++) x[c] = x_in[c];
        for (int l = 0; l < LV; l++) {
            const int idx = ins[l];
            for (int c = 0; c < 6; c++) buf[(l * P + idx) * 6 + c] = x[c];
            for (int lag = 0; lag <= idx; lag++) {
                const int older = (idx - lag + P) % P;
                for (int c = 0; c < 6; c++) corr[(l * P + lag) * 6 + c] += x[c] * buf[(l * P + older) * 6 + c];
                counts[l * P + lag]++;
            }
            ins[l] = (idx + 1) % P;
            for (int c = 0; c < 6; c++) acc[l * 6 + c] += x[c];
            if (++bcount[l] < P) break;
            for (int c = 0; c < 6; c++) {
                x[c] = acc[l * 6 + c] / (double)P;
                acc[l * 6 + c] = 0.0;
            }
            bcount[l] = 0;
        }
    }
};
}  // namespace orc
