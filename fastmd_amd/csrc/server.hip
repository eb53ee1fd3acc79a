// server.hip -- request loop with continuous batching on one GPU.
//
// A worker thread owns the Engine.  Submitted jobs wait in a FIFO admission
// queue until their device footprint fits the HBM budget; an admitted job's
// state is created on the device once and stays resident until it finishes.
// Each turn of the loop launches kGraphSteps steps of the resident set on the
// main stream and, while they run, reads back the jobs that finished in the
// previous launch and creates the newly admitted ones (host staging, uploads
// and their initial list build and forces on a side stream), so the GPU waits
// only for the status check between launches.  Jobs that arrive while others
// run join the next launch.  Each replica stops on the device at its own step
// count, so jobs of different lengths share launches.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "engine.h"
#include "host/admission.h"

namespace fmd {

uint64_t system_bytes(const fmd_system_desc* d);

// Owned deep copy of a system descriptor.
struct DescCopy {
    fmd_system_desc d{};
    std::vector<float> pos, vel, masses, lj, bp, ap, trange, tdata;
    std::vector<int32_t> img, mol, bonds, btypes, angles, tidx, tn;
    explicit DescCopy(const fmd_system_desc& s) : d(s) {
        auto cp = [](auto& v, const auto* p, size_t n) {
            if (p && n) v.assign(p, p + n);
            return v.empty() ? nullptr : v.data();
        };
        const size_t n = s.natoms, nt = s.ntypes;
        d.pos = cp(pos, s.pos, 4 * n);
        d.vel = cp(vel, s.vel, 3 * n);
        d.image = cp(img, s.image, 3 * n);
        d.mol_id = cp(mol, s.mol_id, n);
        d.bonds = cp(bonds, s.bonds, 2 * (size_t)s.nbonds);
        d.bond_types = cp(btypes, s.bond_types, (size_t)s.nbonds);
        d.angles = cp(angles, s.angles, 4 * (size_t)s.nangles);
        d.masses = cp(masses, s.masses, nt);
        d.lj = cp(lj, s.lj, 2 * nt * nt);
        d.bond_params = cp(bp, s.bond_params, 4 * (size_t)s.nbond_types);
        d.angle_params = cp(ap, s.angle_params, 2 * (size_t)s.nangle_types);
        const size_t ntab = s.ntables > 0 ? (size_t)s.ntables : 0;
        size_t npts = 0;
        for (size_t k = 0; ntab && s.table_n && k < ntab; k++) npts += (size_t)std::max(0, s.table_n[k]);
        d.table_idx = ntab ? cp(tidx, s.table_idx, nt * nt) : nullptr;
        d.table_n = cp(tn, s.table_n, ntab);
        d.table_range = cp(trange, s.table_range, 2 * ntab);
        d.table_data = cp(tdata, s.table_data, 3 * npts);
    }
};

struct Job {
    int64_t id;
    std::unique_ptr<DescCopy> desc;
    int64_t nsteps = 0, remaining = 0;
    uint64_t bytes = 0;   // charged against the budget (system_bytes, then the engine's charge)
    int kmult = 1;        // neighbor-capacity multiplier to create with
    int state = FMD_JOB_QUEUED;
    int32_t status = FMD_OK;
    int sys = -1;
    double t_submit = 0, t_admit = 0, t_done = 0;
    fmd_thermo last{};
    std::vector<float> pos, vel;
};

class Server {
  public:
    Server(int device, const fmd_server_opts& o)
        : eng_(device, o.hbm_budget),
          graph_steps_(o.graph_steps > 0 ? o.graph_steps : kGraphSteps),
          aq_(eng_.ok() ? eng_.budget() : 0, o.max_jobs > 0 ? o.max_jobs : 256),
          t0_(std::chrono::steady_clock::now()) {
        if (eng_.ok()) worker_ = std::thread([this] { loop(); });
    }
    ~Server() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        if (worker_.joinable()) worker_.join();
        if (const char* v = std::getenv("FMD_SERVER_PROFILE"))
            if (v[0] == '1')
                std::fprintf(stderr, "[server] overlapped host work %.3f s, enqueue %.3f s, device wait %.3f s, "
                             "between launches %.3f s, requeued %lld\n",
                             prof_[0], prof_[1], prof_[2], prof_[3], (long long)requeued_);
    }
    bool ok() const { return eng_.ok(); }
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

    int32_t submit(const fmd_system_desc* d, int64_t nsteps, int64_t* id) {
        if (!d || nsteps <= 0 || !id) return FMD_ERR_ARG;
        uint64_t bytes = system_bytes(d);
        if (bytes == 0) return FMD_ERR_ARG;
        if (!aq_.fits_ever(bytes)) return FMD_ERR_BUDGET;
        auto j = std::make_unique<Job>();
        j->desc = std::make_unique<DescCopy>(*d);
        j->nsteps = j->remaining = nsteps;
        j->bytes = bytes;
        {
            std::lock_guard<std::mutex> g(m_);
            j->id = next_id_++;
            j->t_submit = now();
            *id = j->id;
            aq_.push(AdmitItem{j->id, bytes});
            jobs_[j->id] = std::move(j);
            pending_++;
        }
        cv_.notify_all();
        return FMD_OK;
    }
    int32_t poll(int64_t id, int32_t* st) {
        std::lock_guard<std::mutex> g(m_);
        auto it = jobs_.find(id);
        if (it == jobs_.end() || !st) return FMD_ERR_ARG;
        *st = it->second->state;
        return FMD_OK;
    }
    int32_t result(int64_t id, fmd_job_result* r, float* pos4, float* vel3) {
        std::lock_guard<std::mutex> g(m_);
        auto it = jobs_.find(id);
        if (it == jobs_.end() || !r) return FMD_ERR_ARG;
        Job& j = *it->second;
        if (j.state != FMD_JOB_DONE && j.state != FMD_JOB_FAILED) return FMD_ERR_STATE;
        r->status = j.status;
        r->natoms = j.desc->d.natoms;
        r->steps = j.nsteps;
        r->t_submit = j.t_submit; r->t_admit = j.t_admit; r->t_done = j.t_done;
        r->last = j.last;
        if (pos4 && !j.pos.empty()) std::memcpy(pos4, j.pos.data(), j.pos.size() * 4);
        if (vel3 && !j.vel.empty()) std::memcpy(vel3, j.vel.data(), j.vel.size() * 4);
        return FMD_OK;
    }
    int32_t drain() {
        std::unique_lock<std::mutex> g(m_);
        done_cv_.wait(g, [this] { return pending_ == 0; });
        return FMD_OK;
    }
    int32_t set_block(int kernel, int cls, int block) {
        if (kernel < 0 || kernel > 2 || cls < 0 || cls >= kNumClasses) return FMD_ERR_ARG;
        if (kernel == 0 && block != 64 && block != 128 && block != 256) return FMD_ERR_ARG;
        if (kernel == 1 && block != kGroupSlots) return FMD_ERR_ARG;  // one workgroup per group of tiles
        if (kernel == 2 && block != 0 && block != 1) return FMD_ERR_ARG;  // neighbor layout
        std::lock_guard<std::mutex> g(m_);
        blocks_.push_back({kernel, cls, block});
        return FMD_OK;
    }
    void stats(double* bs, int64_t* gl, int32_t* res, int32_t* q) {
        std::lock_guard<std::mutex> g(m_);
        if (bs) *bs = bead_steps_;
        if (gl) *gl = launches_;
        if (res) *res = aq_.resident();
        if (q) *q = (int32_t)aq_.queued();
    }

  private:
    void finish(Job& j, int32_t status) {
        // caller holds m_
        j.status = status;
        j.state = status == FMD_OK ? FMD_JOB_DONE : FMD_JOB_FAILED;
        j.t_done = now();
        pending_--;
        done_cv_.notify_all();
    }

    // Reads back a finished job and frees it; the caller guarantees no launch
    // in flight uses it.
    void retire(Job* j, int32_t status) {
        fmd_thermo th{};
        if (eng_.get_thermo(j->sys, &th)) th = fmd_thermo{};
        std::vector<float> pos(4 * (size_t)j->desc->d.natoms), vel(3 * (size_t)j->desc->d.natoms);
        eng_.get_state(j->sys, pos.data(), vel.data(), nullptr, false);
        eng_.destroy(j->sys, false);
        std::lock_guard<std::mutex> g(m_);
        j->last = th;
        j->pos.swap(pos);
        j->vel.swap(vel);
        aq_.release(j->bytes);
        finish(*j, status);
    }

    // Creates an admitted job on the side stream; returns true if it is
    // resident (else it failed, or went back to the head of the queue).
    bool admit_one(Job* j) {
        int32_t sys = -1;
        CreateOpts o;
        o.kmult0 = j->kmult;
        o.side = true;
        int32_t rc = eng_.create(&j->desc->d, &sys, &o);
        if (!rc) rc = eng_.begin_run_side({sys}, j->nsteps);
        std::lock_guard<std::mutex> g(m_);
        if (rc == FMD_ERR_BUDGET && sys < 0 && o.need > j->bytes && aq_.fits_ever(o.need)) {
            // a denser start needs a larger list than admission priced: the job
            // waits at the head of the queue for that much memory
            aq_.release(j->bytes);
            j->bytes = o.need;
            j->kmult = o.kmult;
            aq_.push_front(AdmitItem{j->id, j->bytes});
            requeued_++;
            return false;
        }
        if (rc) {
            aq_.release(j->bytes);
            if (sys >= 0) eng_.destroy(sys, false);
            finish(*j, rc);
            return false;
        }
        if (o.charged != j->bytes) {  // the engine's charge is the one admission tracks
            aq_.adjust((int64_t)o.charged - (int64_t)j->bytes);
            j->bytes = o.charged;
        }
        j->sys = sys;
        j->state = FMD_JOB_RUNNING;
        j->t_admit = now();
        return true;
    }

    void loop() {
        std::vector<Job*> running;            // resident jobs of the next launch
        std::vector<std::pair<Job*, int32_t>> done;  // finished in the last launch
        while (true) {
            std::vector<Job*> admitted;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || !running.empty() || !done.empty() || aq_.queued() > 0; });
                if (stop_) break;
                for (const BlockSet& b : blocks_) eng_.set_tuned(b.kernel, b.cls, b.block);
                blocks_.clear();
            }
            // 1. launch the resident set (asynchronous)
            std::vector<int> ids;
            for (Job* j : running) ids.push_back(j->sys);
            const int64_t k = graph_steps_;
            int32_t rc = FMD_OK;
            const double ts0 = now();
            if (!ids.empty()) {
                // a hipGraph pays off once the resident set is stable: right after
                // an admission or a completion the set is new and likely to change
                // again, so its steps are issued directly instead of captured
                const bool stable = (ids == prev_ids_);
                prev_ids_ = ids;
                rc = eng_.launch_steps(ids, k, stable);
            }
            const double ts1 = now();
            // 2. while it runs: read back and free the jobs of the last launch,
            //    then admit and create new jobs on the side stream
            eng_.set_defer_free(true);
            for (auto& jd : done) retire(jd.first, jd.second);
            done.clear();
            {
                std::lock_guard<std::mutex> g(m_);
                for (const AdmitItem& it : aq_.admit()) admitted.push_back(jobs_[it.job].get());
            }
            std::vector<Job*> created;
            for (Job* j : admitted)
                if (admit_one(j)) created.push_back(j);
            const double ts2 = now();
            // 3. end of the launch
            if (!rc) rc = eng_.sync();
            eng_.flush_deferred();
            const double ts3 = now();
            prof_[0] += ts2 - ts1;
            prof_[1] += ts1 - ts0;
            prof_[2] += ts3 - ts2;
            double bs = 0;
            std::vector<int> failed;
            if (!ids.empty()) {
                for (Job* j : running) bs += (double)j->desc->d.natoms * (double)std::min<int64_t>(k, j->remaining);
                if (!rc) eng_.check_errors(ids, &failed);
            }
            std::vector<Job*> keep;
            for (Job* j : running) {
                j->remaining -= k;
                const bool bad = rc != FMD_OK || std::find(failed.begin(), failed.end(), j->sys) != failed.end();
                if (bad || j->remaining <= 0) {
                    const int32_t eng_err = bad && !rc ? eng_.replica_error(j->sys) : FMD_OK;
                    done.push_back({j, bad ? (rc ? rc : eng_err) : FMD_OK});
                } else {
                    keep.push_back(j);
                }
            }
            for (Job* j : created) keep.push_back(j);
            {
                std::lock_guard<std::mutex> g(m_);
                bead_steps_ += bs;
                if (!ids.empty()) launches_++;
            }
            running.swap(keep);
            // batches of sets that are gone (the device is idle here)
            std::vector<int> rids;
            for (Job* j : running) rids.push_back(j->sys);
            eng_.trim_batches(rids);
            prof_[3] += now() - ts3;
        }
        for (auto& jd : done) retire(jd.first, jd.second);
    }

    struct BlockSet { int kernel, cls, block; };
    Engine eng_;
    int graph_steps_;
    std::vector<BlockSet> blocks_;  // block shapes to apply before the next launch
    std::vector<int> prev_ids_;     // resident set of the previous launch
    double prof_[4] = {0, 0, 0, 0}; // seconds: host work beside a launch, enqueue, device wait, between launches
    int64_t requeued_ = 0;          // jobs handed back to the queue needing more memory
    AdmissionQueue aq_;
    std::chrono::steady_clock::time_point t0_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    std::map<int64_t, std::unique_ptr<Job>> jobs_;
    int64_t next_id_ = 1;
    int64_t pending_ = 0;
    bool stop_ = false;
    double bead_steps_ = 0;
    int64_t launches_ = 0;
    std::thread worker_;
};

}  // namespace fmd

using fmd::Server;

struct fmd_server {
    Server* s;
};

extern "C" {

int32_t fmd_server_create(int32_t device, const fmd_server_opts* o, fmd_server** out) {
    if (!out) return FMD_ERR_ARG;
    fmd_server_opts def{0, 0, 0};
    Server* s = new Server(device, o ? *o : def);
    if (!s->ok()) {
        delete s;
        return FMD_ERR_HIP;
    }
    *out = new fmd_server{s};
    return FMD_OK;
}
int32_t fmd_server_submit(fmd_server* s, const fmd_system_desc* d, int64_t nsteps, int64_t* job) {
    return s ? s->s->submit(d, nsteps, job) : FMD_ERR_ARG;
}
int32_t fmd_server_poll(fmd_server* s, int64_t job, int32_t* state) {
    return s ? s->s->poll(job, state) : FMD_ERR_ARG;
}
int32_t fmd_server_result(fmd_server* s, int64_t job, fmd_job_result* r, float* pos4, float* vel3) {
    return s ? s->s->result(job, r, pos4, vel3) : FMD_ERR_ARG;
}
int32_t fmd_server_drain(fmd_server* s) { return s ? s->s->drain() : FMD_ERR_ARG; }
double fmd_server_now(const fmd_server* s) { return s ? s->s->now() : 0.0; }
int32_t fmd_server_stats(const fmd_server* s, double* bead_steps, int64_t* graph_launches,
                         int32_t* resident, int32_t* queued) {
    if (!s) return FMD_ERR_ARG;
    s->s->stats(bead_steps, graph_launches, resident, queued);
    return FMD_OK;
}
int32_t fmd_server_set_block(fmd_server* s, int32_t kernel, int32_t size_class, int32_t block) {
    return s ? s->s->set_block(kernel, size_class, block) : FMD_ERR_ARG;
}
void fmd_server_destroy(fmd_server* s) {
    if (!s) return;
    delete s->s;
    delete s;
}

}  // extern "C"
