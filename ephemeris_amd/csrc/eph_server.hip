// Serving loop: continuous batching of ensemble and test-particle jobs over
// state resident in HBM (eph_server_* of include/ephemeris_hip.h).
//
// Two loops, one per job kind, each on its own stream (and, under
// eph_server_start, its own host thread), so short particle quanta never wait
// for a long ensemble quantum.  One tick of a loop: admit queued jobs of its
// kind first-in first-out while their footprint fits the memory budget, the
// device's free memory and the slots; re-form the ensemble batch or the
// particle pool when membership changed (resident state is copied device to
// device); advance every resident job of the kind by one quantum; retire
// finished jobs with their results.
//
// Results never depend on co-residents: ensemble kernels never mix systems
// and every job advances in calls of exactly ens_quantum steps from its own
// epoch; a particle is one thread replaying precomputed field records, and
// its state is stored whole between launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "eph_runtime.hpp"

namespace eph {
bool mass_sanity(const eph_mass &m, std::string *why);

namespace {

constexpr int64_t kSysOverhead = 16384;      // plan tables, class tables, per-system metadata

struct SJob {
    int64_t id = 0;
    int kind = EPH_JOB_ENSEMBLE;
    int status = EPH_JOB_QUEUED;
    std::string error;
    int64_t steps_total = 0, steps_done = 0, bytes = 0, units = 0;
    double t_submit = -1, t_admit = -1, t_done = -1;
    // ensemble: inputs, then results in place
    std::vector<int64_t> sys_n;
    std::vector<eph_mass> masses;
    std::vector<eph_geopotential> gps;
    std::vector<eph_ring> rings;
    std::vector<double> t_eph;          // n_sys x {hi, lo}
    // particles: inputs r, v (n x 3); results r, v (n x 6), extrema (n x 2)
    std::vector<double> r, v, ext;
    int64_t col = 0;                    // first pool column while resident
};

double clock_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

bool finite_state(const eph_mass &m) {
    const double x[12] = {m.r.x.hi, m.r.x.lo, m.r.y.hi, m.r.y.lo, m.r.z.hi, m.r.z.lo,
                          m.v.x.hi, m.v.x.lo, m.v.y.hi, m.v.y.lo, m.v.z.hi, m.v.z.lo};
    for (double e : x)
        if (!std::isfinite(e)) return false;
    return std::isfinite(m.w.x) && std::isfinite(m.w.y) && std::isfinite(m.w.z);
}

}  // namespace

struct Server {
    eph_server_config cfg{};
    int device = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    std::mutex mu;                       // jobs, queue, counters
    std::mutex tick_mu[2];               // one tick of each kind at a time
    std::atomic<int> snap_waiting{0};    // snapshot callers waiting for the ensemble tick lock
    std::map<int64_t, std::unique_ptr<SJob>> jobs;
    std::deque<int64_t> queue;
    int64_t next_id = 1;
    int64_t mem_used = 0, mem_peak = 0, mem_budget = 0;
    int64_t n_running = 0;
    int64_t n_run[2] = {0, 0}, n_queued[2] = {0, 0};   // per job kind
    eph_server_stats st{};

    // ensemble residency
    Batch *ens = nullptr;
    std::vector<int64_t> ens_owner;      // job id of every system of `ens`
    std::vector<int64_t> ens_jobs;       // resident ensemble jobs, admission order
    bool ens_dirty = false;

    // particle residency
    JMain *rec = nullptr;                // entry 0: field epoch; entry 1 + 25 s + (k-1): stage k of step s
    int nf = 0;
    hipStream_t pst = nullptr;
    double *P = nullptr;
    int64_t npp = 0;
    PBlock *d_blocks = nullptr;
    int64_t blocks_cap = 0;
    std::vector<int64_t> part_jobs;      // resident particle jobs in pool order
    bool part_dirty = false;

    // worker
    std::thread worker[2];
    std::atomic<bool> stop_flag{false}, worker_on{false};
    std::condition_variable cv_work, cv_idle;
    std::string worker_error;

    ~Server();
    int init(const eph_server_config &c, Batch *field);
    int submit(std::unique_ptr<SJob> j, int64_t *id);
    int tick(int64_t *n_active);
    int tick_kind(int kind);
    int admit(int kind, std::vector<SJob *> &fresh);
    int regroup_ensembles(const std::vector<SJob *> &new_ens);
    int rebuild_pool(const std::vector<SJob *> &new_part);
    int launch_particles_tick(std::vector<PBlock> &host_blocks, int64_t *pairs, int64_t *psteps);
    int retire_ensembles();
    int retire_particles();
    void mark_done(SJob *j);
    bool has_work(int kind) { return n_queued[kind] > 0 || n_run[kind] > 0; }
    void fail_job(SJob *j, const std::string &why);
    // the job map is shared with submitting threads: look up under the lock
    SJob *find_job(int64_t id) {
        std::lock_guard<std::mutex> lk(mu);
        auto it = jobs.find(id);
        return it == jobs.end() ? nullptr : it->second.get();
    }
    PView pool_view() const {
        return PView{P, npp, (int)npp, rec, 1, nf, cfg.part_R < 0 ? cfg.part_R * cfg.part_R : 0.0};
    }
};

Server::~Server() {
    stop_flag = true;
    cv_work.notify_all();
    for (auto &w : worker)
        if (w.joinable()) w.join();
    if (pst) (void)hipStreamSynchronize(pst);
    delete ens;
    if (P) (void)hipFree(P);
    if (rec) (void)hipFree(rec);
    if (d_blocks) (void)hipFree(d_blocks);
    if (pst) (void)hipStreamDestroy(pst);
}

int Server::init(const eph_server_config &c, Batch *field) {
    cfg = c;
    if (!(cfg.ens_dt != 0) || cfg.ens_quantum <= 0 || !(cfg.part_dt != 0) || cfg.part_quantum <= 0 ||
        cfg.part_max_steps < 0 || cfg.mem_budget < 0)
        return fail(EPH_EINVAL, "eph_server_create: bad configuration");
    HIP_CHECK(hipGetDevice(&device));
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    mem_budget = cfg.mem_budget ? cfg.mem_budget : (int64_t)(0.4 * (double)free_b);
    HIP_CHECK(hipStreamCreateWithFlags(&pst, hipStreamNonBlocking));
    if (field) {
        if (field->n_sys != 1 || !field->all_small())
            return fail(EPH_EINVAL, "the particle field must be one system of at most 48 bodies");
        // records of the field's whole horizon, from a private copy of it
        Batch *copy = nullptr;
        int rc = regroup_batch(field, std::vector<int>{0}, 0, nullptr, nullptr, nullptr, 0, nullptr, 0, nullptr, &copy);
        if (rc) return rc;
        std::unique_ptr<Batch> own(copy);
        nf = (int)copy->nb;
        const size_t entries = 1 + 25 * (size_t)cfg.part_max_steps;
        if (hipMalloc(&rec, sizeof(JMain) * kSmallMax * entries) != hipSuccess)
            return fail(EPH_ENOMEM, "field record allocation failed");
        if ((rc = field_epoch_record(copy, rec))) return rc;
        if (cfg.part_max_steps)
            if ((rc = copy->integrate_small(cfg.part_dt, cfg.part_max_steps, rec + kSmallMax, cfg.part_max_steps,
                                            nullptr, nullptr)))
                return rc;
    }
    st.mem_budget = mem_budget;
    return EPH_OK;
}

int Server::submit(std::unique_ptr<SJob> j, int64_t *id) {
    std::lock_guard<std::mutex> lk(mu);
    j->id = next_id++;
    j->t_submit = clock_s(t0);
    if (j->bytes > mem_budget)
        return fail(EPH_ENOMEM, "job needs " + std::to_string(j->bytes) + " bytes, more than the memory budget");
    if (id) *id = j->id;
    queue.push_back(j->id);
    ++n_queued[j->kind];
    jobs[j->id] = std::move(j);
    ++st.jobs_submitted;
    cv_work.notify_all();
    return EPH_OK;
}

void Server::fail_job(SJob *j, const std::string &why) {
    std::lock_guard<std::mutex> lk(mu);
    if (j->status == EPH_JOB_RUNNING) {
        mem_used -= j->bytes;
        --n_running;
        --n_run[j->kind];
    }
    j->status = EPH_JOB_FAILED;
    j->error = why;
    j->t_done = clock_s(t0);
    ++st.jobs_failed;
}

// First in, first out within the kind: stop at the first job of the kind
// that does not fit now.
int Server::admit(int kind, std::vector<SJob *> &fresh) {
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    std::lock_guard<std::mutex> lk(mu);
    const bool is_ens = kind == EPH_JOB_ENSEMBLE;
    int64_t used_slots = 0;                // systems or particles resident
    for (int64_t id : is_ens ? ens_jobs : part_jobs) {
        auto it = jobs.find(id);
        if (it != jobs.end() && it->second->status == EPH_JOB_RUNNING)
            used_slots += is_ens ? (int64_t)it->second->sys_n.size() : it->second->units;
    }
    const int64_t max_slots = is_ens ? cfg.max_systems : cfg.max_particles;
    const int64_t resident = is_ens ? (int64_t)F_COUNT * 8 * (ens ? ens->nb : 0) : (int64_t)P_ALLOC_ROWS * 8 * npp;
    int64_t reserve = 0;                   // a re-formed batch holds old and new copies at once
    for (auto qi = queue.begin(); qi != queue.end();) {
        SJob *j = jobs[*qi].get();
        if (j->kind != kind) { ++qi; continue; }
        const int64_t want = is_ens ? (int64_t)j->sys_n.size() : j->units;
        const bool slots_ok = !max_slots || used_slots + want <= max_slots;
        const bool mem_ok = mem_used + j->bytes <= mem_budget &&
                            (int64_t)free_b >= reserve + resident + 2 * j->bytes;
        if (!slots_ok || !mem_ok) {
            // Failed for good only when it could not be admitted even with no
            // job of either kind resident (the budget was checked at
            // submission); otherwise it waits for jobs to finish.
            const bool never_slots = max_slots && want > max_slots;
            const bool never_mem = n_running == 0 && fresh.empty() && (int64_t)free_b < 2 * j->bytes;
            if (never_slots || never_mem) {
                qi = queue.erase(qi);
                --n_queued[kind];
                j->status = EPH_JOB_FAILED;
                j->error = never_slots ? "job exceeds the server's slots" : "job does not fit in free device memory";
                j->t_done = clock_s(t0);
                ++st.jobs_failed;
                continue;
            }
            break;
        }
        qi = queue.erase(qi);
        --n_queued[kind];
        j->status = EPH_JOB_RUNNING;
        j->t_admit = clock_s(t0);
        mem_used += j->bytes;
        mem_peak = std::max(mem_peak, mem_used);
        reserve += 2 * j->bytes;
        ++n_running;
        ++n_run[kind];
        used_slots += want;
        fresh.push_back(j);
    }
    return EPH_OK;
}

int Server::regroup_ensembles(const std::vector<SJob *> &new_ens) {
    std::vector<int> keep;
    std::vector<int64_t> owner;
    if (ens)
        for (int64_t s = 0; s < ens->n_sys; ++s) {
            SJob *j = find_job(ens_owner[s]);
            if (j && j->status == EPH_JOB_RUNNING) {
                keep.push_back((int)s);
                owner.push_back(ens_owner[s]);
            }
        }
    std::vector<int64_t> sys_n;
    std::vector<eph_mass> masses;
    std::vector<eph_geopotential> gps;
    std::vector<eph_ring> rings;
    std::vector<double> teph;
    for (SJob *j : new_ens) {
        const int64_t gb = (int64_t)gps.size(), rb = (int64_t)rings.size();
        for (eph_mass m : j->masses) {
            if (m.gp >= 0) m.gp += gb;
            if (m.ring >= 0) m.ring += rb;
            masses.push_back(m);
        }
        gps.insert(gps.end(), j->gps.begin(), j->gps.end());
        rings.insert(rings.end(), j->rings.begin(), j->rings.end());
        sys_n.insert(sys_n.end(), j->sys_n.begin(), j->sys_n.end());
        teph.insert(teph.end(), j->t_eph.begin(), j->t_eph.end());
        for (size_t s = 0; s < j->sys_n.size(); ++s) owner.push_back(j->id);
    }
    Batch *nb = nullptr;
    int rc = regroup_batch(ens, keep, (int64_t)sys_n.size(), sys_n.data(), masses.data(), gps.data(),
                           (int64_t)gps.size(), rings.data(), (int64_t)rings.size(), teph.data(), &nb);
    if (rc) return rc;
    delete ens;
    ens = nb;
    ens_owner = owner;
    ens_jobs.clear();
    for (int64_t id : owner)
        if (ens_jobs.empty() || ens_jobs.back() != id) ens_jobs.push_back(id);
    return EPH_OK;
}

int Server::rebuild_pool(const std::vector<SJob *> &new_part) {
    std::vector<SJob *> order;
    for (int64_t id : part_jobs) {
        SJob *j = find_job(id);
        if (j && j->status == EPH_JOB_RUNNING) order.push_back(j);
    }
    order.insert(order.end(), new_part.begin(), new_part.end());
    auto pad = [](int64_t n) { return (n + kPartBS - 1) / kPartBS * kPartBS; };
    int64_t npp_new = 0;
    for (SJob *j : order) npp_new += pad(j->units);
    double *Pn = nullptr;
    if (npp_new) {
        HIP_CHECK(hipMalloc(&Pn, sizeof(double) * P_ALLOC_ROWS * npp_new));
        int64_t c = 0;
        std::vector<std::vector<double>> staging;
        for (SJob *j : order) {
            const int64_t w = pad(j->units);
            if (std::find(new_part.begin(), new_part.end(), j) == new_part.end()) {   // resident: device to device
                HIP_CHECK(hipMemcpy2DAsync(Pn + c, sizeof(double) * npp_new, P + j->col, sizeof(double) * npp,
                                           sizeof(double) * w, P_COUNT_HOST, hipMemcpyDeviceToDevice, pst));
            } else {
                staging.emplace_back((size_t)P_COUNT_HOST * w, 0.0);
                std::vector<double> &h = staging.back();
                for (int64_t i = 0; i < j->units; ++i)
                    for (int q = 0; q < 3; ++q) {
                        h[(size_t)(2 * q) * w + i] = j->r[3 * i + q];
                        h[(size_t)(6 + 2 * q) * w + i] = j->v[3 * i + q];
                    }
                HIP_CHECK(hipMemcpy2DAsync(Pn + c, sizeof(double) * npp_new, h.data(), sizeof(double) * w,
                                           sizeof(double) * w, P_COUNT_HOST, hipMemcpyHostToDevice, pst));
            }
            j->col = c;
            c += w;
        }
        HIP_CHECK(hipStreamSynchronize(pst));
    }
    if (P) HIP_CHECK(hipFree(P));
    P = Pn;
    npp = npp_new;
    part_jobs.clear();
    for (SJob *j : order) part_jobs.push_back(j->id);
    const int64_t nblk = npp / kPartBS;
    if (nblk > blocks_cap) {
        if (d_blocks) HIP_CHECK(hipFree(d_blocks));
        d_blocks = nullptr;
        blocks_cap = std::max<int64_t>(nblk, 2 * blocks_cap);
        HIP_CHECK(hipMalloc(&d_blocks, sizeof(PBlock) * blocks_cap));
    }
    // first acceleration of the new particles from the field's epoch record
    if (!new_part.empty()) {
        std::vector<PBlock> hb(nblk, PBlock{0, 0, 1});
        for (SJob *j : new_part)
            for (int64_t b = 0; b * kPartBS < j->units; ++b)
                hb[j->col / kPartBS + b] = PBlock{0, -1, (int32_t)std::min<int64_t>(kPartBS, j->units - b * kPartBS)};
        HIP_CHECK(hipMemcpyAsync(d_blocks, hb.data(), sizeof(PBlock) * nblk, hipMemcpyHostToDevice, pst));
        launch_particles(pool_view(), 0, 0.0, DtK{}, pst, d_blocks);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(pst));
    }
    return EPH_OK;
}

int Server::launch_particles_tick(std::vector<PBlock> &hb, int64_t *pairs, int64_t *psteps) {
    const int64_t nblk = npp / kPartBS;
    hb.assign(nblk, PBlock{0, 0, 1});
    *pairs = *psteps = 0;
    for (int64_t id : part_jobs) {
        SJob *j = find_job(id);
        const int64_t n = std::min(cfg.part_quantum, j->steps_total - j->steps_done);
        for (int64_t b = 0; b * kPartBS < j->units; ++b)
            hb[j->col / kPartBS + b] = PBlock{1 + 25 * j->steps_done, (int32_t)n,
                                              (int32_t)std::min<int64_t>(kPartBS, j->units - b * kPartBS)};
        *pairs += n * 25 * j->units * nf;
        *psteps += n * j->units;
    }
    if (!nblk) return EPH_OK;
    HIP_CHECK(hipMemcpyAsync(d_blocks, hb.data(), sizeof(PBlock) * nblk, hipMemcpyHostToDevice, pst));
    launch_particles(pool_view(), 0, cfg.part_dt, stage_dtk(), pst, d_blocks);
    HIP_CHECK(hipGetLastError());
    return EPH_OK;
}

void Server::mark_done(SJob *j) {
    std::lock_guard<std::mutex> lk(mu);
    j->status = EPH_JOB_DONE;
    j->t_done = clock_s(t0);
    mem_used -= j->bytes;
    --n_running;
    --n_run[j->kind];
    ++st.jobs_done;
}

// Move finished jobs' results to the host and free their residency.
int Server::retire_ensembles() {
    std::vector<SJob *> done;
    for (int64_t id : ens_jobs) {
        SJob *j = find_job(id);
        if (j->status == EPH_JOB_RUNNING && j->steps_done >= j->steps_total) done.push_back(j);
    }
    if (done.empty()) return EPH_OK;
    std::vector<eph_mass> all(ens->nb);
    int rc = ens->download(all.data());
    if (rc) return rc;
    for (SJob *j : done) {
        int64_t k = 0, s0 = -1;
        for (int64_t s = 0; s < ens->n_sys && s0 < 0; ++s)
            if (ens_owner[s] == j->id) s0 = s;
        bool ok = true;
        for (int64_t s = s0; s < s0 + (int64_t)j->sys_n.size(); ++s) {
            for (int64_t g = ens->sys_off[s]; g < ens->sys_off[s + 1]; ++g, ++k) {
                eph_mass m = all[g];
                m.gp = j->masses[k].gp;               // model indices local to the job
                m.ring = j->masses[k].ring;
                ok = ok && finite_state(m);
                j->masses[k] = m;
            }
            j->t_eph[2 * (s - s0)] = ens->t_eph[s].hi;
            j->t_eph[2 * (s - s0) + 1] = ens->t_eph[s].lo;
        }
        if (!ok) fail_job(j, "state became non-finite");
        else mark_done(j);
    }
    ens_dirty = true;
    return EPH_OK;
}

int Server::retire_particles() {
    for (int64_t id : part_jobs) {
        SJob *j = find_job(id);
        if (j->status != EPH_JOB_RUNNING || j->steps_done < j->steps_total) continue;
        const int64_t n = j->units;
        std::vector<double> h((size_t)P_COUNT_HOST * n);
        HIP_CHECK(hipMemcpy2D(h.data(), sizeof(double) * n, P + j->col, sizeof(double) * npp, sizeof(double) * n,
                              P_COUNT_HOST, hipMemcpyDeviceToHost));
        j->r.assign(6 * n, 0.0);
        j->v.assign(6 * n, 0.0);
        j->ext.assign(2 * n, 0.0);
        bool ok = true;
        for (int64_t i = 0; i < n; ++i) {
            for (int c = 0; c < 6; ++c) {
                j->r[6 * i + c] = h[(size_t)c * n + i];
                j->v[6 * i + c] = h[(size_t)(6 + c) * n + i];
                ok = ok && std::isfinite(j->r[6 * i + c]) && std::isfinite(j->v[6 * i + c]);
            }
            j->ext[2 * i] = h[(size_t)15 * n + i];
            j->ext[2 * i + 1] = h[(size_t)16 * n + i];
        }
        part_dirty = true;
        if (!ok) fail_job(j, "particle state became non-finite");
        else mark_done(j);
    }
    return EPH_OK;
}

// One tick of the loop of `kind`.
int Server::tick_kind(int kind) {
    std::lock_guard<std::mutex> tk(tick_mu[kind]);
    HIP_CHECK(hipSetDevice(device));
    const double t_begin = clock_s(t0);
    std::vector<SJob *> fresh;
    int rc = admit(kind, fresh);
    if (rc) return rc;
    double pairs = 0, sys_steps = 0, psteps = 0;
    if (kind == EPH_JOB_ENSEMBLE) {
        if (!fresh.empty() || ens_dirty) {
            if ((rc = regroup_ensembles(fresh))) {
                const std::string why = std::string("ensemble admission failed: ") + eph_last_error();
                for (SJob *j : fresh) fail_job(j, why);
                return rc;
            }
            ens_dirty = false;
        }
        if (!ens) return EPH_OK;
        if ((rc = ens->integrate(cfg.ens_dt, cfg.ens_quantum, 0))) return rc;
        for (int64_t n : ens->sys_n) pairs += (double)n * (double)(n - 1);
        pairs *= 25.0 * (double)cfg.ens_quantum;
        sys_steps = (double)cfg.ens_quantum * (double)ens->n_sys;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (int64_t id : ens_jobs) jobs.at(id)->steps_done += cfg.ens_quantum;
        }
        if ((rc = retire_ensembles())) return rc;
    } else {
        if (!fresh.empty() || part_dirty) {
            if ((rc = rebuild_pool(fresh))) {
                const std::string why = std::string("particle admission failed: ") + eph_last_error();
                for (SJob *j : fresh) fail_job(j, why);
                return rc;
            }
            part_dirty = false;
        }
        if (!npp) return EPH_OK;
        std::vector<PBlock> hb;
        int64_t pp = 0, ps = 0;
        if ((rc = launch_particles_tick(hb, &pp, &ps))) return rc;
        HIP_CHECK(hipStreamSynchronize(pst));
        pairs = (double)pp;
        psteps = (double)ps;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (int64_t id : part_jobs) {
                SJob *j = jobs.at(id).get();
                j->steps_done += std::min(cfg.part_quantum, j->steps_total - j->steps_done);
            }
        }
        if ((rc = retire_particles())) return rc;
    }
    std::lock_guard<std::mutex> lk(mu);
    st.pair_interactions += pairs;
    st.rk12_system_steps += sys_steps;
    st.particle_steps += psteps;
    ++st.ticks;
    st.busy_s += clock_s(t0) - t_begin;
    if (n_running == 0 && queue.empty()) cv_idle.notify_all();
    return EPH_OK;
}

int Server::tick(int64_t *n_active) {
    int rc = tick_kind(EPH_JOB_ENSEMBLE);
    if (rc) return rc;
    if ((rc = tick_kind(EPH_JOB_PARTICLES))) return rc;
    std::lock_guard<std::mutex> lk(mu);
    if (n_active) *n_active = n_running;
    return EPH_OK;
}

}  // namespace eph

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using eph::Server;
using eph::SJob;
using eph::fail;

struct eph_system;
static Server *S(eph_server *s) { return reinterpret_cast<Server *>(s); }

extern "C" {

int eph_server_create(const eph_server_config *cfg, eph_system *field, eph_server **out) {
    if (!cfg || !out) return fail(EPH_EINVAL, "eph_server_create: null argument");
    auto srv = std::make_unique<Server>();
    int rc = srv->init(*cfg, reinterpret_cast<eph::Batch *>(field));
    if (rc) return rc;
    *out = reinterpret_cast<eph_server *>(srv.release());
    return EPH_OK;
}

int eph_server_submit_ensemble(eph_server *srv, int64_t n_sys, const int64_t *sys_n, const eph_mass *masses,
                               const eph_geopotential *gps, int64_t n_gp, const eph_ring *rings, int64_t n_ring,
                               const double *t_eph, int64_t n_steps, int64_t *job) {
    if (!srv || n_sys <= 0 || !sys_n || !masses || n_gp < 0 || n_ring < 0 || (n_gp && !gps) || (n_ring && !rings))
        return fail(EPH_EINVAL, "eph_server_submit_ensemble: bad argument");
    Server *s = S(srv);
    if (n_steps <= 0 || n_steps % s->cfg.ens_quantum)
        return fail(EPH_EINVAL, "ensemble span must be a positive multiple of the ensemble quantum");
    auto j = std::make_unique<SJob>();
    j->kind = EPH_JOB_ENSEMBLE;
    j->steps_total = n_steps;
    int64_t nb = 0;
    for (int64_t k = 0; k < n_sys; ++k) {
        if (sys_n[k] <= 0 || sys_n[k] > (1 << 20)) return fail(EPH_EINVAL, "system size out of range");
        nb += sys_n[k];
    }
    j->sys_n.assign(sys_n, sys_n + n_sys);
    j->masses.assign(masses, masses + nb);
    if (n_gp) j->gps.assign(gps, gps + n_gp);
    if (n_ring) j->rings.assign(rings, rings + n_ring);
    j->t_eph.assign(2 * n_sys, 0.0);
    if (t_eph) std::copy(t_eph, t_eph + 2 * n_sys, j->t_eph.begin());
    for (const auto &g : j->gps)
        if (g.Nz < 2 || g.Nz > EPH_GP_MAX_DEGREE || g.Nt < 0 || g.Nt > g.Nz)
            return fail(EPH_EINVAL, "bad geopotential degree");
    for (const auto &r : j->rings)
        if (r.N < 0 || r.N > EPH_RING_MAX_DISKS) return fail(EPH_EINVAL, "bad ring disk count");
    std::string why;
    for (int64_t g = 0; g < nb; ++g) {
        const eph_mass &m = j->masses[g];
        if (m.gp >= n_gp || m.ring >= n_ring) return fail(EPH_EINVAL, "model index out of range");
        std::string w;
        if (!eph::mass_sanity(m, &w) || !eph::finite_state(m)) {
            char id[9] = {0};
            memcpy(id, &m.sid, 8);
            why += std::string("<") + id + ">: " + (w.empty() ? "non-finite state; " : w);
        }
    }
    if (!why.empty()) return fail(EPH_ESANITY, "job rejected, sanity check failed: " + why);
    j->units = nb;
    j->bytes = (int64_t)eph::F_COUNT * 8 * (nb + 64) + eph::kSysOverhead * n_sys;
    return s->submit(std::move(j), job);
}

int eph_server_submit_particles(eph_server *srv, int64_t n, const double *r, const double *v, int64_t n_steps,
                                int64_t *job) {
    if (!srv || n <= 0 || !r || !v) return fail(EPH_EINVAL, "eph_server_submit_particles: bad argument");
    Server *s = S(srv);
    if (!s->rec) return fail(EPH_EINVAL, "the server has no particle field");
    if (n_steps <= 0 || n_steps > s->cfg.part_max_steps)
        return fail(EPH_EINVAL, "particle span must be within 1..part_max_steps");
    for (int64_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(r[i]) || !std::isfinite(v[i])) return fail(EPH_ESANITY, "job rejected: non-finite particle state");
    auto j = std::make_unique<SJob>();
    j->kind = EPH_JOB_PARTICLES;
    j->steps_total = n_steps;
    j->units = n;
    j->r.assign(r, r + 3 * n);
    j->v.assign(v, v + 3 * n);
    j->bytes = (int64_t)eph::P_ALLOC_ROWS * 8 * ((n + eph::kPartBS - 1) / eph::kPartBS * eph::kPartBS);
    return s->submit(std::move(j), job);
}

int eph_server_tick(eph_server *srv, int64_t *n_active) {
    if (!srv) return fail(EPH_EINVAL, "null server");
    Server *s = S(srv);
    if (s->worker_on) return fail(EPH_EBUSY, "the server's worker thread is running");
    return s->tick(n_active);
}

int eph_server_start(eph_server *srv) {
    if (!srv) return fail(EPH_EINVAL, "null server");
    Server *s = S(srv);
    if (s->worker_on) return EPH_OK;
    s->stop_flag = false;
    s->worker_on = true;
    for (int kind = 0; kind < 2; ++kind)
        s->worker[kind] = std::thread([s, kind] {
            if (hipSetDevice(s->device) != hipSuccess) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->worker_error = "hipSetDevice failed in a worker";
                s->cv_idle.notify_all();
                return;
            }
            while (!s->stop_flag) {
                {
                    std::unique_lock<std::mutex> lk(s->mu);
                    s->cv_work.wait_for(lk, std::chrono::milliseconds(50),
                                        [s, kind] { return s->stop_flag || s->has_work(kind); });
                    if (s->stop_flag) break;
                    if (!s->has_work(kind)) continue;
                }
                // a snapshot caller waiting for the tick lock goes first
                while (kind == EPH_JOB_ENSEMBLE && s->snap_waiting.load() > 0 && !s->stop_flag) std::this_thread::yield();
                if (s->tick_kind(kind)) {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->worker_error = eph_last_error();
                    s->stop_flag = true;
                    s->cv_idle.notify_all();
                    s->cv_work.notify_all();
                    break;
                }
            }
        });
    return EPH_OK;
}

int eph_server_wait(eph_server *srv, double timeout_s) {
    if (!srv) return fail(EPH_EINVAL, "null server");
    Server *s = S(srv);
    std::unique_lock<std::mutex> lk(s->mu);
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    while (!s->worker_error.size() && !(s->queue.empty() && s->n_running == 0)) {
        if (s->cv_idle.wait_until(lk, deadline) == std::cv_status::timeout) {
            if (s->queue.empty() && s->n_running == 0) break;
            return fail(EPH_EBUSY, "eph_server_wait: jobs still pending at the time limit");
        }
    }
    if (!s->worker_error.empty()) return fail(EPH_EDEVICE, "server worker stopped: " + s->worker_error);
    return EPH_OK;
}

int eph_server_stop(eph_server *srv) {
    if (!srv) return fail(EPH_EINVAL, "null server");
    Server *s = S(srv);
    s->stop_flag = true;
    s->cv_work.notify_all();
    for (auto &w : s->worker)
        if (w.joinable()) w.join();
    s->worker_on = false;
    return EPH_OK;
}

int eph_server_job_info(eph_server *srv, int64_t job, eph_job_info *info) {
    if (!srv || !info) return fail(EPH_EINVAL, "null argument");
    Server *s = S(srv);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return fail(EPH_EINVAL, "unknown job");
    const SJob &j = *it->second;
    memset(info, 0, sizeof *info);
    info->kind = j.kind;
    info->status = j.status;
    info->steps_total = j.steps_total;
    info->steps_done = j.steps_done;
    info->bytes = j.bytes;
    info->units = j.units;
    info->t_submit = j.t_submit;
    info->t_admit = j.t_admit;
    info->t_done = j.t_done;
    strncpy(info->error, j.error.c_str(), sizeof info->error - 1);
    return EPH_OK;
}

static SJob *done_job(Server *s, int64_t job, int kind) {
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) { fail(EPH_EINVAL, "unknown job"); return nullptr; }
    SJob *j = it->second.get();
    if (j->kind != kind) { fail(EPH_EINVAL, "job is of the other kind"); return nullptr; }
    if (j->status != EPH_JOB_DONE) { fail(EPH_EBUSY, "job has no result (not done)"); return nullptr; }
    return j;
}

int eph_server_result_ensemble(eph_server *srv, int64_t job, eph_mass *masses, double *t_eph) {
    if (!srv) return fail(EPH_EINVAL, "null server");
    Server *s = S(srv);
    std::lock_guard<std::mutex> lk(s->mu);
    SJob *j = done_job(s, job, EPH_JOB_ENSEMBLE);
    if (!j) return EPH_EINVAL;
    if (masses) std::copy(j->masses.begin(), j->masses.end(), masses);
    if (t_eph) std::copy(j->t_eph.begin(), j->t_eph.end(), t_eph);
    return EPH_OK;
}

int eph_server_snapshot_ensemble(eph_server *srv, int64_t job, eph_mass *masses, double *t_eph, int64_t *steps_done) {
    if (!srv) return fail(EPH_EINVAL, "null server");
    Server *s = S(srv);
    // between ticks of the ensemble loop the resident batch holds whole
    // quanta; the worker lets a waiting snapshot take the lock between ticks
    s->snap_waiting.fetch_add(1);
    std::unique_lock<std::mutex> tk(s->tick_mu[EPH_JOB_ENSEMBLE]);
    s->snap_waiting.fetch_sub(1);
    SJob *j = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(job);
        if (it == s->jobs.end()) return fail(EPH_EINVAL, "unknown job id");
        j = it->second.get();
        if (j->kind != EPH_JOB_ENSEMBLE) return fail(EPH_EINVAL, "job is of the other kind");
        if (j->status == EPH_JOB_FAILED) return fail(EPH_EINVAL, "job failed: " + j->error);
        if (j->status != EPH_JOB_RUNNING) {               // queued: its input; done: its result
            if (masses) std::copy(j->masses.begin(), j->masses.end(), masses);
            if (t_eph) std::copy(j->t_eph.begin(), j->t_eph.end(), t_eph);
            if (steps_done) *steps_done = j->steps_done;
            return EPH_OK;
        }
    }
    // running: its systems in the resident batch
    HIP_CHECK(hipSetDevice(s->device));
    std::vector<eph_mass> all((size_t)s->ens->nb);
    int rc = s->ens->download(all.data());
    if (rc) return rc;
    int64_t s0 = -1;
    for (int64_t q = 0; q < s->ens->n_sys && s0 < 0; ++q)
        if (s->ens_owner[(size_t)q] == j->id) s0 = q;
    if (s0 < 0) return fail(EPH_EDEVICE, "running job not found in the resident batch");
    int64_t k = 0;
    for (int64_t q = s0; q < s0 + (int64_t)j->sys_n.size(); ++q) {
        for (int64_t g = s->ens->sys_off[(size_t)q]; g < s->ens->sys_off[(size_t)q + 1]; ++g, ++k) {
            if (!masses) continue;
            eph_mass m = all[(size_t)g];
            m.gp = j->masses[(size_t)k].gp;                // model indices local to the job
            m.ring = j->masses[(size_t)k].ring;
            masses[k] = m;
        }
        if (t_eph) {
            t_eph[2 * (q - s0)] = s->ens->t_eph[(size_t)q].hi;
            t_eph[2 * (q - s0) + 1] = s->ens->t_eph[(size_t)q].lo;
        }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (steps_done) *steps_done = j->steps_done;
    return EPH_OK;
}

int eph_server_result_particles(eph_server *srv, int64_t job, double *r, double *v, double *extrema) {
    if (!srv) return fail(EPH_EINVAL, "null server");
    Server *s = S(srv);
    std::lock_guard<std::mutex> lk(s->mu);
    SJob *j = done_job(s, job, EPH_JOB_PARTICLES);
    if (!j) return EPH_EINVAL;
    if (r) std::copy(j->r.begin(), j->r.end(), r);
    if (v) std::copy(j->v.begin(), j->v.end(), v);
    if (extrema) std::copy(j->ext.begin(), j->ext.end(), extrema);
    return EPH_OK;
}

int eph_server_release(eph_server *srv, int64_t job) {
    if (!srv) return fail(EPH_EINVAL, "null server");
    Server *s = S(srv);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return fail(EPH_EINVAL, "unknown job");
    if (it->second->status != EPH_JOB_DONE && it->second->status != EPH_JOB_FAILED)
        return fail(EPH_EBUSY, "job is still queued or running");
    s->jobs.erase(it);
    return EPH_OK;
}

int eph_server_stats_read(eph_server *srv, eph_server_stats *out) {
    if (!srv || !out) return fail(EPH_EINVAL, "null argument");
    Server *s = S(srv);
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->jobs_queued = (int64_t)s->queue.size();
    out->jobs_running = s->n_running;
    out->mem_used = s->mem_used;
    out->mem_peak = s->mem_peak;
    out->mem_budget = s->mem_budget;
    out->t_now = eph::clock_s(s->t0);
    return EPH_OK;
}

int eph_server_destroy(eph_server *srv) {
    if (!srv) return EPH_OK;
    delete S(srv);
    return EPH_OK;
}

}  // extern "C"
