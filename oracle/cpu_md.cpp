// cpu_md.cpp -- TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times
// beside the GPU path (cpu_baseline.kind = "port").  An OpenMP restatement of the
// reference's Langevin step (src/main.cu:403-630): BAOAB (src/integrate/langevin.cu:34-130),
// skin-triggered Verlet rebuild with the forced 20-step rebuild (src/main.cu:400-528),
// cell binning and a full list (src/neighbor/verlet_list.cu:58-335), LJ
// (src/force/lj.cu:27-63), FENE+WCA (src/force/fene.cu:21-65) and harmonic angle
// (src/force/angle.cu:22-79), FP32 arithmetic.  Each atom gathers its own forces
// (no atomics), parallel over atoms with OpenMP.
#include <omp.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "oracle.hpp"

using namespace orc;

namespace {

struct Topo {
    std::vector<int> boff, blst, btyp;        // per atom: bond partners
    std::vector<int> aoff, alst;              // per atom: angle index list
};

inline float mi(float d, float L, float invL) { return d - L * std::rint(d * invL); }

}  // namespace

extern "C" {

// Runs nsteps Langevin steps in place; returns the number of neighbor-list
// rebuilds.  pos4/vel3/img3 are updated.  threads <= 0 uses omp default.
long orc_cpu_md_langevin(int n, float* pos4, float* vel3, int* img3, const float* masses,
                         float eps, float sig, float rc, float skin, float L, float dt, float T,
                         float Tdamp, uint64_t seed, int nbonds, const int* bonds, const float* bp,
                         int nang, const int* ang4, const float* ap, long nsteps, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    const float invL = 1.0f / L, rc2 = rc * rc, rcs = rc + skin, rcs2 = rcs * rcs;
    const float half_dt = 0.5f * dt;
    const float c1 = std::exp(-(1.0f / Tdamp) * dt), c2 = std::sqrt(1.0f - c1 * c1);
    Topo tp;
    tp.boff.assign(n + 1, 0);
    for (int b = 0; b < nbonds; b++) { tp.boff[bonds[2 * b] + 1]++; tp.boff[bonds[2 * b + 1] + 1]++; }
    for (int i = 0; i < n; i++) tp.boff[i + 1] += tp.boff[i];
    tp.blst.resize(tp.boff[n]);
    {
        std::vector<int> f(tp.boff.begin(), tp.boff.end() - 1);
        for (int b = 0; b < nbonds; b++) {
            tp.blst[f[bonds[2 * b]]++] = bonds[2 * b + 1];
            tp.blst[f[bonds[2 * b + 1]]++] = bonds[2 * b];
        }
    }
    tp.aoff.assign(n + 1, 0);
    for (int a = 0; a < nang; a++)
        for (int q = 0; q < 3; q++) tp.aoff[ang4[4 * a + q] + 1]++;
    for (int i = 0; i < n; i++) tp.aoff[i + 1] += tp.aoff[i];
    tp.alst.resize(tp.aoff[n]);
    {
        std::vector<int> f(tp.aoff.begin(), tp.aoff.end() - 1);
        for (int a = 0; a < nang; a++)
            for (int q = 0; q < 3; q++) tp.alst[f[ang4[4 * a + q]]++] = a;
    }
    const float k_b = bp[0], R0 = bp[1], beps = bp[2], bsig = bp[3];
    const float k_a = ap[0], th0 = ap[1];

    CellGrid g = cell_grid(L, rcs);
    const int nc = g.nx * g.nx * g.nx;
    const int maxn = 256;
    std::vector<int> nbr((size_t)n * maxn), nn(n), cstart(nc + 1), citems(n), cid(n);
    std::vector<float> ref(3 * (size_t)n), f3(3 * (size_t)n, 0.0f);

    auto rebuild = [&]() {
        std::fill(cstart.begin(), cstart.end(), 0);
        for (int i = 0; i < n; i++) { cid[i] = cell_id(pos4 + 4 * i, g); cstart[cid[i] + 1]++; }
        for (int c = 0; c < nc; c++) cstart[c + 1] += cstart[c];
        std::vector<int> fill(cstart.begin(), cstart.end() - 1);
        for (int i = 0; i < n; i++) citems[fill[cid[i]]++] = i;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            int c[3];
            cell_coords(pos4 + 4 * i, g, c);
            int cnt = 0, visited[27], nv = 0;
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        int x = (c[0] + dx + g.nx) % g.nx, y = (c[1] + dy + g.nx) % g.nx,
                            z = (c[2] + dz + g.nx) % g.nx;
                        int cc = (z * g.nx + y) * g.nx + x;
                        bool seen = false;
                        for (int v = 0; v < nv; v++) seen |= visited[v] == cc;
                        if (seen) continue;
                        visited[nv++] = cc;
                        for (int t = cstart[cc]; t < cstart[cc + 1]; t++) {
                            int j = citems[t];
                            if (j == i) continue;
                            float ddx = mi(pos4[4 * i] - pos4[4 * j], L, invL);
                            float ddy = mi(pos4[4 * i + 1] - pos4[4 * j + 1], L, invL);
                            float ddz = mi(pos4[4 * i + 2] - pos4[4 * j + 2], L, invL);
                            if (ddx * ddx + ddy * ddy + ddz * ddz >= rcs2) continue;
                            bool ex = false;
                            for (int e = tp.boff[i]; e < tp.boff[i + 1]; e++) ex |= tp.blst[e] == j;
                            if (!ex && cnt < maxn) nbr[(size_t)i * maxn + cnt++] = j;
                        }
                    }
            nn[i] = cnt;
            ref[3 * i] = pos4[4 * i]; ref[3 * i + 1] = pos4[4 * i + 1]; ref[3 * i + 2] = pos4[4 * i + 2];
        }
    };
    auto forces = [&]() {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            const float* pi = pos4 + 4 * i;
            float fx = 0, fy = 0, fz = 0;
            const int* nl = &nbr[(size_t)i * maxn];
            const float s2 = sig * sig;
            for (int k = 0; k < nn[i]; k++) {
                const float* pj = pos4 + 4 * nl[k];
                float dx = mi(pi[0] - pj[0], L, invL), dy = mi(pi[1] - pj[1], L, invL),
                      dz = mi(pi[2] - pj[2], L, invL);
                float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < rc2) {
                    float ri = 1.0f / r2, sr2 = s2 * ri, sr6 = sr2 * sr2 * sr2;
                    float fr = 24.0f * eps * ri * (2.0f * sr6 * sr6 - sr6);
                    fx += fr * dx; fy += fr * dy; fz += fr * dz;
                }
            }
            for (int e = tp.boff[i]; e < tp.boff[i + 1]; e++) {
                const float* pj = pos4 + 4 * tp.blst[e];
                float dx = mi(pi[0] - pj[0], L, invL), dy = mi(pi[1] - pj[1], L, invL),
                      dz = mi(pi[2] - pj[2], L, invL);
                float r2 = dx * dx + dy * dy + dz * dz, R02 = R0 * R0;
                r2 = std::fmin(std::fmax(r2, 1e-6f), R02 * 0.9999f);
                float ff = -k_b * R02 / (R02 - r2);
                if (r2 < bsig * bsig * 1.2599210498948732f) {
                    float ri = 1.0f / r2, sr2 = bsig * bsig * ri, sr6 = sr2 * sr2 * sr2;
                    ff += 24.0f * beps * ri * (2.0f * sr6 * sr6 - sr6);
                }
                fx += ff * dx; fy += ff * dy; fz += ff * dz;
            }
            for (int e = tp.aoff[i]; e < tp.aoff[i + 1]; e++) {
                const int* A = ang4 + 4 * tp.alst[e];
                const float *a0 = pos4 + 4 * A[0], *a1 = pos4 + 4 * A[1], *a2 = pos4 + 4 * A[2];
                float ax = mi(a0[0] - a1[0], L, invL), ay = mi(a0[1] - a1[1], L, invL),
                      az = mi(a0[2] - a1[2], L, invL);
                float bx = mi(a2[0] - a1[0], L, invL), by = mi(a2[1] - a1[1], L, invL),
                      bz = mi(a2[2] - a1[2], L, invL);
                float ra2 = ax * ax + ay * ay + az * az, rb2 = bx * bx + by * by + bz * bz;
                float ria = 1.0f / std::sqrt(ra2), rib = 1.0f / std::sqrt(rb2);
                float c = (ax * bx + ay * by + az * bz) * ria * rib;
                c = std::fmin(std::fmax(c, -1.0f), 1.0f);
                float s = std::fmax(std::sqrt(1.0f - c * c), 1e-6f);
                float pref = 2.0f * k_a * (std::acos(c) - th0) / s;
                float fix = pref * (bx * ria * rib - c * ax * ria * ria);
                float fiy = pref * (by * ria * rib - c * ay * ria * ria);
                float fiz = pref * (bz * ria * rib - c * az * ria * ria);
                float fkx = pref * (ax * ria * rib - c * bx * rib * rib);
                float fky = pref * (ay * ria * rib - c * by * rib * rib);
                float fkz = pref * (az * ria * rib - c * bz * rib * rib);
                if (A[0] == i) { fx += fix; fy += fiy; fz += fiz; }
                else if (A[2] == i) { fx += fkx; fy += fky; fz += fkz; }
                else { fx -= fix + fkx; fy -= fiy + fky; fz -= fiz + fkz; }
            }
            f3[3 * i] = fx; f3[3 * i + 1] = fy; f3[3 * i + 2] = fz;
        }
    };
    rebuild();
    forces();
    long nreb = 1;
    int since = 0;
    for (long step = 1; step <= nsteps; step++) {
        float maxd = 0.0f;
#pragma omp parallel for schedule(static) reduction(max : maxd)
        for (int i = 0; i < n; i++) {
            float* r = pos4 + 4 * i;
            float* v = vel3 + 3 * i;
            float m = masses[type_of(r[3])];
            float hdtm = half_dt / m;
            float z[3];
            normal3(seed, (uint32_t)i, step, z);
            float ns = c2 * std::sqrt(T / m);
            for (int d = 0; d < 3; d++) {
                v[d] += hdtm * f3[3 * i + d];
                r[d] += half_dt * v[d];
                v[d] = c1 * v[d] + ns * z[d];
                r[d] += half_dt * v[d];
                int ix = (int)std::floor(r[d] * invL);
                r[d] -= (float)ix * L;
                img3[3 * i + d] += ix;
            }
            float dx = r[0] - ref[3 * i], dy = r[1] - ref[3 * i + 1], dz = r[2] - ref[3 * i + 2];
            dx -= L * std::round(dx * invL); dy -= L * std::round(dy * invL); dz -= L * std::round(dz * invL);
            float d2 = dx * dx + dy * dy + dz * dz;
            maxd = d2 > maxd ? d2 : maxd;
        }
        since++;
        if (maxd > 0.25f * skin * skin || since >= 20) { rebuild(); nreb++; since = 0; }
        forces();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            float hdtm = half_dt / masses[type_of(pos4[4 * i + 3])];
            for (int d = 0; d < 3; d++) vel3[3 * i + d] += hdtm * f3[3 * i + d];
        }
    }
    return nreb;
}

int orc_cpu_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
