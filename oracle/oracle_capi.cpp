// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY: extern "C" entry points of the
// CPU oracle for tests/ (ctypes), __graft_entry__.smoke() and bench.py's
// cpu_baseline leg.  Never linked into the product library.
#include <cstdio>
#include "oracle.hpp"

using namespace orc;

extern "C" {

float orc_min_image(float dr, float L, float invL) { return min_image_ref(dr, L, invL); }
float orc_min_image_fixed(float a, float b, float L) { return min_image_fixed_f(a, b, L); }
// n pairs of points (xyz each): out[k] = 1 if the device's grid-unit LJ test keeps pair k
void orc_lj_inside_grid(int n, const float* a, const float* b, float rc2, float L, int32_t* out) {
    for (int k = 0; k < n; k++) {
        const float pa[3] = {a[3 * k], a[3 * k + 1], a[3 * k + 2]};
        const float pb[3] = {b[3 * k], b[3 * k + 1], b[3 * k + 2]};
        out[k] = lj_inside_grid(pa, pb, rc2, L) ? 1 : 0;
    }
}
float orc_wrap(float x, float L, float invL) { return wrap_ref(x, L, invL); }
uint32_t orc_morton3d(float x, float y, float z, float invL) { return morton3d(x, y, z, invL); }

void orc_morton_perm(const float* pos4, int n, float invL, int* perm) {
    std::vector<int> p = morton_perm(pos4, n, invL);
    for (int i = 0; i < n; i++) perm[i] = p[i];
}

// Reference cell ids (src/neighbor/verlet_list.cu:58-75); returns nx.
int orc_cell_ids(const float* pos4, int n, float L, float rcs, int* cell) {
    CellGrid g = cell_grid(L, rcs);
    for (int i = 0; i < n; i++) cell[i] = cell_id(pos4 + 4 * i, g);
    return g.nx;
}

int orc_build_exclusions(int n, const int* bonds, int nbonds, int* off, int* lst, int cap) {
    std::vector<int> o, l;
    build_exclusions(n, bonds, nbonds, o, l);
    for (int i = 0; i <= n; i++) off[i] = o[i];
    int m = (int)l.size();
    for (int e = 0; e < m && e < cap; e++) lst[e] = l[e];
    return m;
}

// Ordered pair set; method 0 = brute force, 1 = cell grid.  Writes up to cap
// pairs (2 ints each), returns the total number of pairs.
long orc_pair_set(const float* pos4, int n, float L, float rcs, const int* eoff, const int* elst,
                  int method, int* ij, long cap) {
    std::vector<int> out;
    float invL = 1.0f / L;
    if (method == 0)
        pair_set_brute(pos4, n, L, invL, rcs, eoff, elst, out);
    else
        pair_set_cells(pos4, n, L, invL, rcs, eoff, elst, out);
    long np = (long)out.size() / 2;
    for (long p = 0; p < np && p < cap; p++) { ij[2 * p] = out[2 * p]; ij[2 * p + 1] = out[2 * p + 1]; }
    return np;
}

void orc_lj_literal(const float* pos4, int n, const float* lj2, int ntypes, float rc, float L,
                    const int* eoff, const int* elst, float* f4, float* vir) {
    lj_literal_f32(pos4, n, lj2, ntypes, rc, L, 1.0f / L, eoff, elst, f4, vir);
}
void orc_fene_literal(const float* pos4, int n, const int* bonds, const int* btype, int nb,
                      const float* bp, float L, float* f4, float* vir) {
    fene_literal_f32(pos4, n, bonds, btype, nb, bp, L, 1.0f / L, f4, vir);
}
void orc_angle_literal(const float* pos4, int n, const int* ang4, int na, const float* ap,
                       float L, float* f4, float* vir) {
    angle_literal_f32(pos4, n, ang4, na, ap, L, 1.0f / L, f4, vir);
}

void orc_forces_f64(const float* pos4, int n, float L, int ntypes, const float* lj2, float rc,
                    const float* bp, const float* ap, int clamp_lj, int clamp_fene, float fcap,
                    int terms, const int* eoff, const int* elst, const int* bonds,
                    const int* btype, int nbonds, const int* ang4, int nang, double* f8,
                    double* vir) {
    ForceParams P{ntypes, lj2, rc, bp, ap, clamp_lj, clamp_fene, fcap};
    forces_f64(pos4, n, L, P, terms, eoff, elst, bonds, btype, nbonds, ang4, nang, f8, vir);
}

// forces_f64 with tabulated pairs
void orc_forces_tab_f64(const float* pos4, int n, float L, int ntypes, const float* lj2, float rc,
                        const float* bp, const float* ap, int clamp_lj, int clamp_fene, float fcap,
                        int terms, const int* eoff, const int* elst, const int* bonds,
                        const int* btype, int nbonds, const int* ang4, int nang, int ntables,
                        const int* tab_idx, const int* tab_n, const float* tab_range,
                        const float* tab_data, double* f8, double* vir) {
    ForceParams P{ntypes, lj2, rc, bp, ap, clamp_lj, clamp_fene, fcap};
    P.ntables = ntables; P.tab_idx = tab_idx; P.tab_n = tab_n;
    P.tab_range = tab_range; P.tab_data = tab_data;
    forces_f64(pos4, n, L, P, terms, eoff, elst, bonds, btype, nbonds, ang4, nang, f8, vir);
}

void orc_table_eval(int n, float rmin, float rmax, const float* rows, int nr, const double* r,
                    double* f, double* e) {
    for (int k = 0; k < nr; k++) table_eval_f64(n, rmin, rmax, rows, r[k], f + k, e + k);
}

void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    philox4x32_10(ctr, key, out);
}
void orc_normal3(uint64_t seed, uint32_t gid, int64_t step, float* z) { normal3(seed, gid, step, z); }

// One NH chain application.  prec 0 = FP32 literal (nh_expf), 1 = FP64.
// state: xi[10], vxi[10] in/out (double storage either way).  Returns scale.
double orc_nh_chain(int prec, int M, double Q1, double Qr, double* xi, double* vxi, double KE,
                    double kT, int natoms, double half_dt) {
    if (prec == 0) {
        NHChain<float> c;
        c.M = M; c.Q1 = (float)Q1; c.Qr = (float)Qr;
        for (int k = 0; k < 10; k++) { c.xi[k] = (float)xi[k]; c.vxi[k] = (float)vxi[k]; }
        float s = nh_chain_step<float>(c, (float)KE, (float)kT, natoms, (float)half_dt);
        for (int k = 0; k < 10; k++) { xi[k] = c.xi[k]; vxi[k] = c.vxi[k]; }
        return s;
    }
    NHChain<double> c;
    c.M = M; c.Q1 = Q1; c.Qr = Qr;
    for (int k = 0; k < 10; k++) { c.xi[k] = xi[k]; c.vxi[k] = vxi[k]; }
    double s = nh_chain_step<double>(c, KE, kT, natoms, half_dt);
    for (int k = 0; k < 10; k++) { xi[k] = c.xi[k]; vxi[k] = c.vxi[k]; }
    return s;
}

// Chain energy from a state: form 0 = the reference's conserved-quantity
// chain terms (tests/test_nose_hoover.cu:19-40), 1 = the form the chain's
// dynamics conserves (nh_chain_energy_dyn).
double orc_nh_chain_energy(int form, int M, double Q1, double Qr, const double* xi, const double* vxi,
                           double kT, int natoms) {
    NHChain<double> c;
    c.M = M; c.Q1 = Q1; c.Qr = Qr;
    for (int k = 0; k < 10; k++) { c.xi[k] = xi[k]; c.vxi[k] = vxi[k]; }
    return form == 0 ? nh_chain_energy<double>(c, kT, natoms) : nh_chain_energy_dyn<double>(c, kT, natoms);
}

// ---------------------------------------------------------------------------
// MD driver in the reference's operator order (src/main.cu:403-630).
// Positions/velocities in FP32 (as on the device), forces from forces_f64
// rounded to FP32, KE/virial reductions and NH/barostat scalars in FP64.
// Langevin noise from normal3(seed, gid, step).  Thermo rows (every step):
// step, KE/N, PE/N, T, Pxx, Pxy, Pxz, Pyy, Pyz, Pzz, L  (11 doubles).
// ---------------------------------------------------------------------------
struct OrcSys {
    int n, ntypes;
    float* pos4;   // in/out
    float* vel3;   // in/out
    int* img3;     // in/out
    const float* masses;
    const float* lj2;
    int nbonds;
    const int* bonds;
    const int* btype;
    const float* bp;
    int nang;
    const int* ang4;
    const float* ap;
    float L, rc, skin, dt;
    int ensemble;  // 0 Langevin, 1 NVT-NH, 2 NPT-NH
    float T_start, T_stop, Tdamp, P_start, P_stop, Pdamp;
    int chain;
    uint64_t seed;
    long nsteps_total;
    int clamp_lj, clamp_fene;
    float fcap;
};

static void eval_forces(const OrcSys& s, float L, const std::vector<int>& eoff,
                        const std::vector<int>& elst, std::vector<float>& f4, double* vir,
                        double* pe) {
    ForceParams P{s.ntypes, s.lj2, s.rc, s.bp, s.ap, s.clamp_lj, s.clamp_fene, s.fcap};
    std::vector<double> f8(4 * (size_t)s.n);
    forces_f64(s.pos4, s.n, L, P, 7, eoff.data(), elst.data(), s.bonds, s.btype, s.nbonds,
               s.ang4, s.nang, f8.data(), vir);
    *pe = 0.0;
    for (int i = 0; i < s.n; i++) {
        for (int d = 0; d < 4; d++) f4[4 * i + d] = (float)f8[4 * i + d];
        *pe += f8[4 * i + 3];
    }
}

static void kin_tensor(const OrcSys& s, double* k6) {
    for (int c = 0; c < 6; c++) k6[c] = 0.0;
    for (int i = 0; i < s.n; i++) {
        double m = s.masses[type_of(s.pos4[4 * i + 3])];
        double vx = s.vel3[3 * i], vy = s.vel3[3 * i + 1], vz = s.vel3[3 * i + 2];
        k6[0] += m * vx * vx; k6[1] += m * vx * vy; k6[2] += m * vx * vz;
        k6[3] += m * vy * vy; k6[4] += m * vy * vz; k6[5] += m * vz * vz;
    }
}

static void wrap_atom(float* r, int* img, float L, float invL) {
    for (int d = 0; d < 3; d++) {
        int ix = (int)std::floor(r[d] * invL);
        r[d] -= (float)ix * L;
        img[d] += ix;
    }
}

// Extended-system record (ext, nullable): kExt doubles per row, row 0 before
// the first step and row k after step k:
//   0 K (kinetic energy, sum m v^2 / 2), 1 U (potential energy), 2 V of the box
//   the positions are in, 3 V0*exp(3 eps) (the barostat's volume), 4 Q1, 5 Qr,
//   6 W, 7 eps, 8 v_eps, 9 M, 10..19 xi[k], 20..29 v_xi[k], 30 target kT of the
//   step, 31 target P of the step, 32 the LJ energy at the cutoff summed over the
//   pairs inside it (U - col 32 is the energy of the cut-and-shifted potential,
//   which has the same forces and no jump when a pair crosses rc).
constexpr int kExt = 40;

// Sum over unordered non-excluded pairs inside the LJ cutoff (the grid-unit test
// of the force path) of 4 eps ((sig/rc)^12 - (sig/rc)^6) of the pair's types.
static double lj_cut_energy_raw(const float* pos4, int n, int ntypes, const float* lj2, float rc, float L,
                                const int* eoff, const int* elst) {
    double sum = 0.0;
    const float rc2 = rc * rc;
    for (int i = 0; i < n; i++) {
        const int ti = type_of(pos4[4 * i + 3]);
        if (ti < 0) continue;
        for (int j = i + 1; j < n; j++) {
            const int tj = type_of(pos4[4 * j + 3]);
            if (tj < 0 || excluded(eoff, elst, i, j)) continue;
            if (!lj_inside_grid(pos4 + 4 * i, pos4 + 4 * j, rc2, L)) continue;
            const double eps = lj2[2 * (ti * ntypes + tj)], sig = lj2[2 * (ti * ntypes + tj) + 1];
            const double sr6 = std::pow(sig / (double)rc, 6.0);
            sum += 4.0 * eps * (sr6 * sr6 - sr6);
        }
    }
    return sum;
}
static double lj_cut_energy(const OrcSys& s, float L, const std::vector<int>& eoff,
                            const std::vector<int>& elst) {
    return lj_cut_energy_raw(s.pos4, s.n, s.ntypes, s.lj2, s.rc, L, eoff.data(), elst.data());
}

// The same sum for a given state (exclusions from the bond list).
double orc_lj_cut_energy(const float* pos4, int n, int ntypes, const float* lj2, float rc, float L,
                         const int* bonds, int nbonds) {
    std::vector<int> eoff, elst;
    build_exclusions(n, bonds, nbonds, eoff, elst);
    return lj_cut_energy_raw(pos4, n, ntypes, lj2, rc, L, eoff.data(), elst.data());
}

int orc_md_run(OrcSys* sp, long nsteps, double* thermo, double* ext) {
    OrcSys& s = *sp;
    const int n = s.n;
    std::vector<int> eoff, elst;
    build_exclusions(n, s.bonds, s.nbonds, eoff, elst);
    std::vector<float> f4(4 * (size_t)n);
    double vir[6], pe;
    float L = s.L;
    eval_forces(s, L, eoff, elst, f4, vir, &pe);
    const float half_dt = 0.5f * s.dt;
    const double hdt_d = 0.5 * (double)s.dt, dt_d = (double)s.dt;
    // Langevin constants (src/integrate/langevin.cu:15-21)
    float c1 = std::exp(-(1.0f / s.Tdamp) * s.dt);
    float c2 = std::sqrt(1.0f - c1 * c1);
    // NH init (src/integrate/nose_hoover.cu:309-346), FP64
    NHChain<double> ch;
    ch.M = s.chain > 10 ? 10 : s.chain;
    double Nf = 3.0 * n, kT0 = s.T_start;
    ch.Q1 = Nf * kT0 * (double)s.Tdamp * s.Tdamp;
    ch.Qr = kT0 * (double)s.Tdamp * s.Tdamp;
    for (int k = 0; k < 10; k++) ch.xi[k] = ch.vxi[k] = 0.0;
    Baro<double> baro{0.0, 0.0, (Nf + 3.0) * kT0 * (double)s.Pdamp * s.Pdamp,
                      (double)s.L * s.L * s.L, (double)s.P_start};
    double k6[6];
    kin_tensor(s, k6);
    double carry = 0.5 * (k6[0] + k6[3] + k6[5]);
    auto record = [&](long row, double K, double U, float Lpos, double kT, double Pt) {
        if (!ext) return;
        double* x = ext + (size_t)kExt * row;
        x[0] = K; x[1] = U; x[2] = (double)Lpos * Lpos * Lpos;
        x[3] = baro.V0 * std::exp(3.0 * baro.eps);
        x[4] = ch.Q1; x[5] = ch.Qr; x[6] = baro.W; x[7] = baro.eps; x[8] = baro.veps; x[9] = ch.M;
        for (int k = 0; k < 10; k++) { x[10 + k] = ch.xi[k]; x[20 + k] = ch.vxi[k]; }
        x[30] = kT; x[31] = Pt;
        x[32] = lj_cut_energy(s, Lpos, eoff, elst);
    };
    record(0, carry, pe, L, s.T_start, s.P_start);
    for (long step = 1; step <= nsteps; step++) {
        float frac = (float)(step - 1) / (float)s.nsteps_total;
        float kT = s.T_start + (s.T_stop - s.T_start) * frac;
        float Pt_step = s.P_start + (s.P_stop - s.P_start) * frac;
        float Lpos = L;  // box the positions of this step are placed in
        float invL = 1.0f / L;
        if (s.ensemble == 0) {
            for (int i = 0; i < n; i++) {
                float* r = s.pos4 + 4 * i;
                float* v = s.vel3 + 3 * i;
                float m = s.masses[type_of(r[3])];
                float hdtm = half_dt / m;
                for (int d = 0; d < 3; d++) v[d] += hdtm * f4[4 * i + d];
                for (int d = 0; d < 3; d++) r[d] += half_dt * v[d];
                float z[3];
                normal3(s.seed, (uint32_t)i, step, z);
                float ns = c2 * std::sqrt(kT / m);
                for (int d = 0; d < 3; d++) v[d] = c1 * v[d] + ns * z[d];
                for (int d = 0; d < 3; d++) r[d] += half_dt * v[d];
                wrap_atom(r, s.img3 + 3 * i, L, invL);
            }
        } else {
            float Pt = s.P_start + (s.P_stop - s.P_start) * frac;
            baro.Ptarget = Pt;
            double nh, bs = 1.0, exp_vW = 1.0, fpos = s.dt;
            if (s.ensemble == 1) {
                nh = nh_chain_step<double>(ch, carry, (double)kT, n, hdt_d);
            } else {
                BaroOut<double> o = npt_chain_baro<double>(ch, baro, carry, 0.0, false, (double)kT,
                                                           n, hdt_d, dt_d);
                nh = o.nh_scale;
                bs = o.baro_scale;
                exp_vW = o.exp_vW;
                double h = 0.5 * o.veps_W_dt;
                fpos = dt_d * std::exp(h) * sinch(h);
                L = (float)o.L;
                invL = 1.0f / L;
                Lpos = L;
            }
            float sc = (float)(nh * bs);
            float ew = (float)exp_vW, fp = (float)fpos;
            for (int i = 0; i < n; i++) {
                float* r = s.pos4 + 4 * i;
                float* v = s.vel3 + 3 * i;
                float m = s.masses[type_of(r[3])];
                float hdtm = half_dt / m;
                for (int d = 0; d < 3; d++) v[d] *= sc;
                for (int d = 0; d < 3; d++) v[d] += hdtm * f4[4 * i + d];
                if (s.ensemble == 1)
                    for (int d = 0; d < 3; d++) r[d] += s.dt * v[d];
                else
                    for (int d = 0; d < 3; d++) r[d] = r[d] * ew + v[d] * fp;
                wrap_atom(r, s.img3 + 3 * i, L, invL);
            }
        }
        eval_forces(s, L, eoff, elst, f4, vir, &pe);
        for (int i = 0; i < n; i++) {
            float m = s.masses[type_of(s.pos4[4 * i + 3])];
            float hdtm = half_dt / m;
            for (int d = 0; d < 3; d++) s.vel3[3 * i + d] += hdtm * f4[4 * i + d];
        }
        kin_tensor(s, k6);
        double ke = 0.5 * (k6[0] + k6[3] + k6[5]);
        double vscale = 1.0;
        if (s.ensemble == 1) {
            double sc = nh_chain_step<double>(ch, ke, (double)kT, n, hdt_d);
            carry = ke * sc * sc;
            vscale = sc;
        } else if (s.ensemble == 2) {
            BaroOut<double> o = npt_chain_baro<double>(ch, baro, ke, vir[0] + vir[3] + vir[5], true,
                                                       (double)kT, n, hdt_d, dt_d);
            carry = o.KE_scaled * o.nh_scale * o.nh_scale;
            vscale = o.baro_scale * o.nh_scale;
            L = (float)o.L;
        }
        if (vscale != 1.0) {
            float b = (float)(vscale);
            for (int i = 0; i < 3 * n; i++) s.vel3[i] *= b;
            kin_tensor(s, k6);
        }
        if (thermo) {
            double* t = thermo + 11 * (step - 1);
            double V = (double)L * L * L;
            double kin = 0.5 * (k6[0] + k6[3] + k6[5]);
            t[0] = (double)step;
            t[1] = kin / n;
            t[2] = pe / n;
            t[3] = 2.0 * (kin / n) / 3.0;
            for (int c = 0; c < 6; c++) t[4 + c] = (k6[c] + vir[c]) / V;
            t[10] = L;
        }
        record(step, 0.5 * (k6[0] + k6[3] + k6[5]), pe, Lpos, (double)kT, (double)Pt_step);
    }
    s.L = L;
    return 0;
}

// The literal CPU path, timed on one core as the reference's own CPU
// baseline: every step the brute-force FP32 forces of reference/cpu_reference.hpp
// (lj_literal_f32 over all pairs, fene_literal_f32, angle_literal_f32), summed
// into one force array, and the serial BAOAB Langevin step of
// src/integrate/langevin.cu:34-130 with the normal3 noise.  Runs nsteps steps in
// place; one type of mass masses[0] per type as in OrcSys.
void orc_literal_md_langevin(int n, float* pos4, float* vel3, int* img3, const float* masses,
                             const float* lj2, int ntypes, float rc, float L, float dt, float T,
                             float Tdamp, uint64_t seed, int nbonds, const int* bonds, const int* btype,
                             const float* bp, int nang, const int* ang4, const float* ap, long nsteps) {
    std::vector<int> eoff, elst;
    build_exclusions(n, bonds, nbonds, eoff, elst);
    const float invL = 1.0f / L, half_dt = 0.5f * dt;
    const float c1 = std::exp(-(1.0f / Tdamp) * dt), c2 = std::sqrt(1.0f - c1 * c1);
    std::vector<float> f(4 * (size_t)n), fb(4 * (size_t)n), fa(4 * (size_t)n);
    float vir[6];
    auto forces = [&]() {
        lj_literal_f32(pos4, n, lj2, ntypes, rc, L, invL, eoff.data(), elst.data(), f.data(), vir);
        if (nbonds) {
            fene_literal_f32(pos4, n, bonds, btype, nbonds, bp, L, invL, fb.data(), vir);
            for (size_t q = 0; q < f.size(); q++) f[q] += fb[q];
        }
        if (nang) {
            angle_literal_f32(pos4, n, ang4, nang, ap, L, invL, fa.data(), vir);
            for (size_t q = 0; q < f.size(); q++) f[q] += fa[q];
        }
    };
    forces();
    for (long step = 1; step <= nsteps; step++) {
        for (int i = 0; i < n; i++) {
            float* r = pos4 + 4 * i;
            float* v = vel3 + 3 * i;
            const float m = masses[type_of(r[3])];
            const float hdtm = half_dt / m;
            for (int d = 0; d < 3; d++) v[d] += hdtm * f[4 * i + d];
            for (int d = 0; d < 3; d++) r[d] += half_dt * v[d];
            float z[3];
            normal3(seed, (uint32_t)i, step, z);
            const float ns = c2 * std::sqrt(T / m);
            for (int d = 0; d < 3; d++) v[d] = c1 * v[d] + ns * z[d];
            for (int d = 0; d < 3; d++) r[d] += half_dt * v[d];
            wrap_atom(r, img3 + 3 * i, L, invL);
        }
        forces();
        for (int i = 0; i < n; i++) {
            const float hdtm = half_dt / masses[type_of(pos4[4 * i + 3])];
            for (int d = 0; d < 3; d++) vel3[3 * i + d] += hdtm * f[4 * i + d];
        }
    }
}

// Thermo from state (src/analysis/thermo.cu:140-201): KE/N, PE/N, T=2KE/(3N),
// P_ab = (sum m v_a v_b + W_ab)/V, virial order xx,xy,xz,yy,yz,zz.
void orc_thermo(int n, const float* pos4, const float* vel3, const float* masses, const double* pe_atom,
                const double* vir, float L, double* out) {
    double k6[6] = {0, 0, 0, 0, 0, 0}, pe = 0;
    for (int i = 0; i < n; i++) {
        double m = masses[type_of(pos4[4 * i + 3])];
        double vx = vel3[3 * i], vy = vel3[3 * i + 1], vz = vel3[3 * i + 2];
        k6[0] += m * vx * vx; k6[1] += m * vx * vy; k6[2] += m * vx * vz;
        k6[3] += m * vy * vy; k6[4] += m * vy * vz; k6[5] += m * vz * vz;
        pe += pe_atom[i];
    }
    double ke = 0.5 * (k6[0] + k6[3] + k6[5]);
    double V = (double)L * L * L;
    out[0] = ke / n;
    out[1] = pe / n;
    out[2] = 2.0 * (ke / n) / 3.0;
    for (int c = 0; c < 6; c++) out[3 + c] = (k6[c] + vir[c]) / V;
}

double orc_rg(int n, const float* pos4, const int32_t* image, const int32_t* mol, float L) {
    return rg_ensemble(n, pos4, image, mol, L);
}

// Pushes nsamples samples (6 each) through a fresh correlator; corr[16*16*6]
// and counts[16*16] receive its sums.
void orc_corr_run(int nsamples, const double* samples, double* corr, int32_t* counts) {
    MultiTau m;
    for (int k = 0; k < nsamples; k++) m.push(samples + 6 * (size_t)k);
    for (size_t i = 0; i < m.corr.size(); i++) corr[i] = m.corr[i];
    for (size_t i = 0; i < m.counts.size(); i++) counts[i] = m.counts[i];
}

}  // extern "C"
