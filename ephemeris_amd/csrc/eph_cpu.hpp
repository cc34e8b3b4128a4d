// Host (CPU-thread) back end of the force evaluation and the RK12 stage
// engine: the reference's parallel_option 0 / -1 paths
// (src/physics/mass.cpp:216-397, src/integrators/RungeKutta.impl:107-194)
// on an array of eph_mass, with the pair terms of eph_terms.hpp -- the same
// source the gfx950 kernels compile -- and a pool of worker threads standing
// in for the reference's thread pool (src/utils/threadpool.cpp:161-218).
// Results do not depend on the number of threads: every sum keeps the
// reference's order, and parallel work items write disjoint data.
#pragma once
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/ephemeris_hip.h"
#include "eph_math.hpp"

namespace eph {

// Persistent worker threads; run(n, fn) calls fn(0..n-1), each index once,
// on the workers and the calling thread, and returns when all are done.
class WorkerPool {
public:
    explicit WorkerPool(int n_threads);
    ~WorkerPool();
    WorkerPool(const WorkerPool &) = delete;
    WorkerPool &operator=(const WorkerPool &) = delete;
    void run(int64_t n, const std::function<void(int64_t)> &fn);
    int threads() const { return (int)workers.size() + 1; }

private:
    void work();
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_go, cv_done;
    const std::function<void(int64_t)> *job = nullptr;
    int64_t job_n = 0, next = 0, done = 0;
    uint64_t generation = 0;
    bool quit = false;
};

// One system as the CPU back end sees it.
struct CpuSystem {
    eph_mass *m = nullptr;
    int64_t n = 0;
    const eph_geopotential *gps = nullptr;
    const eph_ring *rings = nullptr;
    // tidal-matrix correction of a combined-integration subsystem
    // (mass.cpp:376-387; none when tidal_children is empty)
    int64_t tidal_parent = -1;
    m3 tidal_matrix{};
    const int64_t *tidal_children = nullptr;
    int64_t n_tidal_children = 0;
};

// mass::deform_by (mass.cpp:78-103) of body i
void cpu_deform_by(const CpuSystem &s, int64_t i);
// msystem::deform (mass.cpp:104-110)
void cpu_deform(const CpuSystem &s, WorkerPool *pool);
// msystem::accel(parallel_option) for parallel_option 0 and -1
// (mass.cpp:216-397): the thread-pool order when pool != nullptr and
// min(N, ceil(N^2/2304)) > 1 (never for parallel_option -1), else the
// single-thread order with action-reaction updates
void cpu_accel(const CpuSystem &s, int parallel_option, WorkerPool *pool);
// msystem::update (mass.cpp:28-48)
void cpu_update(const CpuSystem &s, double t);
// msystem::integrate(dt, n_step, 0) (mass.cpp:399-442) with the CPU RK12
// (RungeKutta.impl:107-194); t_eph advances by dt * n_step
void cpu_integrate(const CpuSystem &s, dd &t_eph, double dt, int64_t n_step, WorkerPool *pool);

}  // namespace eph
