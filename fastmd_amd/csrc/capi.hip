// capi.hip -- extern "C" implementation of include/fmd/fmd.h.
#include <cstring>
#include <string>

#include "engine.h"
#include "host/admission.h"
#include "host/topo.h"

namespace fmd {
uint64_t system_bytes(const fmd_system_desc* d);
}

struct fmd_engine {
    fmd::Engine* e;
};
struct fmd_topo {
    fmd::Topo t;
};

using fmd::Engine;

extern "C" {

const char* fmd_version(void) { return "fastmd-mi355x 0.1 (gfx950)"; }
const char* fmd_last_host_error(void) { return fmd::host_error(); }

int32_t fmd_topo_load_conf(const char* run_conf, fmd_topo** out) {
    if (!run_conf || !out) return FMD_ERR_ARG;
    auto* t = new fmd_topo();
    int32_t rc = fmd::parse_conf(run_conf, t->t);
    if (rc) { delete t; return rc; }
    *out = t;
    return FMD_OK;
}

int32_t fmd_topo_load_data(const char* data_file, fmd_topo** out) {
    if (!data_file || !out) return FMD_ERR_ARG;
    auto* t = new fmd_topo();
    t->t.extra.restart_freq = -1;
    int32_t rc = fmd::parse_lammps_data(data_file, t->t);
    if (rc) { delete t; return rc; }
    t->t.ntypes = (int)t->t.masses.size();
    *out = t;
    return FMD_OK;
}

int32_t fmd_topo_kg_melt(int32_t nchains, int32_t chain_len, double rho, uint64_t gen_seed,
                         fmd_topo** out) {
    if (!out) return FMD_ERR_ARG;
    auto* t = new fmd_topo();
    int32_t rc = fmd::kg_melt(nchains, chain_len, rho, gen_seed, t->t);
    if (rc) { delete t; return rc; }
    *out = t;
    return FMD_OK;
}

int32_t fmd_topo_desc(fmd_topo* t, fmd_system_desc* d) {
    if (!t || !d) return FMD_ERR_ARG;
    fmd::fill_desc(t->t, *d);
    return FMD_OK;
}

int32_t fmd_topo_conf_extra(const fmd_topo* t, fmd_conf_extra* out) {
    if (!t || !out) return FMD_ERR_ARG;
    *out = t->t.extra;
    return FMD_OK;
}

int64_t fmd_topo_exclusions(const fmd_topo* t, int32_t* off, int32_t* lst, int64_t cap) {
    if (!t) return -1;
    std::vector<int32_t> o, l;
    fmd::build_exclusions(t->t.natoms(), t->t.bonds, o, l);
    if (off) std::memcpy(off, o.data(), o.size() * 4);
    if (lst)
        for (int64_t k = 0; k < (int64_t)l.size() && k < cap; k++) lst[k] = l[k];
    return (int64_t)l.size();
}

int32_t fmd_topo_write_data(const fmd_topo* t, const char* path, const float* pos4,
                            const float* vel3, const int32_t* image3, float L, int64_t step) {
    if (!t || !path) return FMD_ERR_ARG;
    return fmd::write_lammps_data(t->t, path, pos4 ? pos4 : t->t.pos4.data(),
                                  vel3 ? vel3 : t->t.vel3.data(), image3 ? image3 : t->t.image3.data(),
                                  L > 0.0f ? L : t->t.L, step);
}

void fmd_topo_free(fmd_topo* t) { delete t; }

// ----- engine -----
int32_t fmd_engine_create(int32_t device, uint64_t hbm_budget, fmd_engine** out) {
    if (!out) return FMD_ERR_ARG;
    Engine* e = new Engine(device, hbm_budget);
    if (!e->ok()) {
        fmd::set_host_error(e->last_error());
        delete e;
        return FMD_ERR_HIP;
    }
    *out = new fmd_engine{e};
    return FMD_OK;
}
void fmd_engine_destroy(fmd_engine* e) {
    if (!e) return;
    delete e->e;
    delete e;
}
const char* fmd_last_error(const fmd_engine* e) { return e ? e->e->last_error() : fmd::host_error(); }
uint64_t fmd_system_bytes(const fmd_system_desc* d) { return fmd::system_bytes(d); }
uint64_t fmd_engine_bytes_in_use(const fmd_engine* e) { return e ? e->e->bytes_in_use() : 0; }

int32_t fmd_system_create(fmd_engine* e, const fmd_system_desc* d, int32_t* sys) {
    if (!e || !sys) return FMD_ERR_ARG;
    return e->e->create(d, sys);
}
int32_t fmd_system_destroy(fmd_engine* e, int32_t sys) { return e ? e->e->destroy(sys) : FMD_ERR_ARG; }

int32_t fmd_system_set_force_cap(fmd_engine* e, int32_t sys, float force_cap) {
    return e ? e->e->set_force_cap(sys, force_cap) : FMD_ERR_ARG;
}
int32_t fmd_system_set_langevin(fmd_engine* e, int32_t sys, float T_start, float T_stop, float Tdamp) {
    return e ? e->e->set_langevin(sys, T_start, T_stop, Tdamp) : FMD_ERR_ARG;
}
int32_t fmd_build_neighbors(fmd_engine* e, int32_t sys) { return e ? e->e->build_neighbors(sys) : FMD_ERR_ARG; }
int32_t fmd_get_pairs(fmd_engine* e, int32_t sys, int64_t* npairs, int32_t* ij, int64_t cap) {
    return e ? e->e->get_pairs(sys, npairs, ij, cap) : FMD_ERR_ARG;
}
int32_t fmd_get_cell_ids(fmd_engine* e, int32_t sys, int32_t* cell_ids, int32_t* nx) {
    return e ? e->e->get_cell_ids(sys, cell_ids, nx) : FMD_ERR_ARG;
}
int32_t fmd_get_morton_keys(fmd_engine* e, int32_t sys, uint32_t* keys) {
    return e ? e->e->get_morton(sys, keys) : FMD_ERR_ARG;
}
int32_t fmd_compute_forces(fmd_engine* e, int32_t sys, uint32_t terms, float* force4,
                           double* virial6, double* pe) {
    return e ? e->e->compute_forces(sys, terms, force4, virial6, pe) : FMD_ERR_ARG;
}

int32_t fmd_step(fmd_engine* e, const int32_t* sys, int32_t nsys, int64_t nsteps) {
    return e ? e->e->step(sys, nsys, nsteps, true) : FMD_ERR_ARG;
}
int32_t fmd_step_plain(fmd_engine* e, int32_t sys, int64_t nsteps) {
    return e ? e->e->step(&sys, 1, nsteps, false) : FMD_ERR_ARG;
}
int64_t fmd_system_step(fmd_engine* e, int32_t sys) { return e ? e->e->step_count(sys) : -1; }
int32_t fmd_system_error(fmd_engine* e, int32_t sys) { return e ? e->e->replica_error(sys) : FMD_ERR_ARG; }
int32_t fmd_system_natoms(fmd_engine* e, int32_t sys) {
    fmd::Replica* r = e ? e->e->rep(sys) : nullptr;
    return r ? r->natoms : -1;
}
int32_t fmd_get_state(fmd_engine* e, int32_t sys, float* pos4, float* vel3, int32_t* image3) {
    return e ? e->e->get_state(sys, pos4, vel3, image3) : FMD_ERR_ARG;
}
int32_t fmd_get_box(fmd_engine* e, int32_t sys, double* L) { return e ? e->e->get_box(sys, L) : FMD_ERR_ARG; }
int32_t fmd_snapshot_size(fmd_engine* e, int32_t sys, uint64_t* bytes) {
    return e ? e->e->snapshot_size(sys, bytes) : FMD_ERR_ARG;
}
int32_t fmd_snapshot(fmd_engine* e, int32_t sys, void* buf, uint64_t cap) {
    return e ? e->e->snapshot(sys, buf, cap) : FMD_ERR_ARG;
}
int32_t fmd_system_resume(fmd_engine* e, const void* buf, uint64_t size, int32_t* sys) {
    return e ? e->e->resume(buf, size, sys) : FMD_ERR_ARG;
}
int32_t fmd_corr_push(fmd_engine* e, int32_t sys, const double* sample6) {
    return e ? e->e->corr_push(sys, sample6) : FMD_ERR_ARG;
}
int32_t fmd_corr_reset(fmd_engine* e, int32_t sys) {
    return e ? e->e->corr_reset(sys) : FMD_ERR_ARG;
}
int32_t fmd_get_corr(fmd_engine* e, int32_t sys, double* corr, int32_t* counts, int32_t* levels) {
    return e ? e->e->get_corr(sys, corr, counts, levels) : FMD_ERR_ARG;
}
int32_t fmd_get_rg(fmd_engine* e, int32_t sys, double* rg, double* per_chain, int32_t cap, int32_t* nchains) {
    return e ? e->e->get_rg(sys, rg, per_chain, cap, nchains) : FMD_ERR_ARG;
}
int32_t fmd_get_rg_noalloc(fmd_engine* e, int32_t sys, double* rg, double* per_chain, int32_t cap,
                           int32_t* nchains) {
    return e ? e->e->get_rg_noalloc(sys, rg, per_chain, cap, nchains) : FMD_ERR_ARG;
}
int32_t fmd_get_state_box(fmd_engine* e, int32_t sys, double* L) {
    return e ? e->e->get_box(sys, L, true) : FMD_ERR_ARG;
}
int32_t fmd_get_thermo(fmd_engine* e, int32_t sys, fmd_thermo* out) {
    return e ? e->e->get_thermo(sys, out) : FMD_ERR_ARG;
}
int32_t fmd_get_thermostat(fmd_engine* e, int32_t sys, fmd_thermostat* out) {
    return e ? e->e->get_thermostat(sys, out) : FMD_ERR_ARG;
}
int32_t fmd_nh_chain_run(fmd_engine* e, const fmd_chain_run* c, double* xi, double* vxi, double* scales) {
    return e ? e->e->nh_chain_run(c, xi, vxi, scales) : FMD_ERR_ARG;
}
int32_t fmd_drain_thermo(fmd_engine* e, int32_t sys, fmd_thermo* out, int32_t cap, int32_t* n,
                         int64_t* lost) {
    return e ? e->e->drain_thermo(sys, out, cap, n, lost) : FMD_ERR_ARG;
}
int64_t fmd_system_rebuilds(fmd_engine* e, int32_t sys) { return e ? e->e->rebuilds(sys) : -1; }

int32_t fmd_autotune(fmd_engine* e, const int32_t* sys, int32_t nsys, int32_t reps) {
    return e ? e->e->autotune(sys, nsys, reps) : FMD_ERR_ARG;
}
int32_t fmd_time_force(fmd_engine* e, const int32_t* sys, int32_t nsys, uint32_t terms, int32_t reps,
                       double* avg_ms, double* alg_bytes, int64_t* pairs) {
    return e ? e->e->time_force(sys, nsys, terms, reps, avg_ms, alg_bytes, pairs) : FMD_ERR_ARG;
}
int32_t fmd_tuned_block(const fmd_engine* e, int32_t kernel, int32_t size_class) {
    if (!e || kernel < 0 || kernel > 2 || size_class < 0 || size_class >= fmd::kNumClasses) return -1;
    return e->e->tuned(kernel, size_class);
}
int32_t fmd_engine_set_tuned(fmd_engine* e, int32_t kernel, int32_t size_class, int32_t value) {
    return e ? e->e->set_tuned(kernel, size_class, value) : FMD_ERR_ARG;
}
int32_t fmd_autotune_times(const fmd_engine* e, int32_t size_class, int32_t* layout, int32_t* block, double* ms,
                           int32_t cap) {
    return e ? e->e->tune_times(size_class, layout, block, ms, cap) : -1;
}
int32_t fmd_size_class(int32_t natoms) { return fmd::size_class(natoms); }

int32_t fmd_admission_replay(int32_t njobs, const double* t_arrive, const uint64_t* bytes,
                             const double* duration, uint64_t budget, int32_t max_jobs,
                             double* t_admit, double* t_finish) {
    if (njobs < 0 || (njobs > 0 && (!t_arrive || !bytes || !duration || !t_admit || !t_finish)))
        return FMD_ERR_ARG;
    for (int k = 1; k < njobs; k++)
        if (t_arrive[k] < t_arrive[k - 1]) return FMD_ERR_ARG;  // arrivals must be sorted
    fmd::AdmissionQueue q(budget, max_jobs > 0 ? max_jobs : 256);
    for (int k = 0; k < njobs; k++) {
        t_admit[k] = -1.0;
        t_finish[k] = -1.0;
        if (!q.fits_ever(bytes[k])) continue;  // rejected at submission
    }
    // event-driven replay: arrivals and completions in time order
    std::vector<std::pair<double, int>> running;  // (finish time, job)
    int next = 0;
    double now = 0.0;
    while (true) {
        // admit everything that can start now
        for (const fmd::AdmitItem& it : q.admit()) {
            t_admit[it.job] = now;
            t_finish[it.job] = now + duration[it.job];
            running.emplace_back(t_finish[it.job], (int)it.job);
        }
        double t_next_arr = 1e300, t_next_fin = 1e300;
        while (next < njobs && !q.fits_ever(bytes[next])) next++;
        if (next < njobs) t_next_arr = t_arrive[next];
        int fin = -1;
        for (size_t r = 0; r < running.size(); r++)
            if (running[r].first < t_next_fin) { t_next_fin = running[r].first; fin = (int)r; }
        if (next >= njobs && fin < 0) break;
        if (t_next_fin <= t_next_arr) {
            now = t_next_fin;
            q.release(bytes[running[fin].second]);
            running.erase(running.begin() + fin);
        } else {
            now = t_next_arr;
            q.push(fmd::AdmitItem{next, bytes[next]});
            next++;
        }
    }
    return FMD_OK;
}

}  // extern "C"
