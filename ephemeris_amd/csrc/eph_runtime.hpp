// Host-side representation of a resident batch (internal to the library).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "eph_kernels.hpp"

namespace eph {

int fail(int code, const std::string &msg);

// return the library status of a failed HIP call from the enclosing function
#define HIP_CHECK(expr)                                                          \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess)                                                    \
            return fail(EPH_EDEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int kNumKernelClasses = 7;
constexpr int P_COUNT_HOST = 17;     // particle fields (PField)
constexpr int P_STAGE_ROWS = kPartScratchRows;   // per-step scratch rows after them (k_particles)
constexpr int P_ALLOC_ROWS = P_COUNT_HOST + P_STAGE_ROWS;
extern const char *const kKernelNames[kNumKernelClasses];

// Per-kernel-class device time from HIP events recorded on the batch stream.
struct Profiler {
    struct Rec { int cls; hipEvent_t a, b; };
    bool enabled = false;
    unsigned mask = ~0u;             // kernel classes timed while enabled
    std::vector<Rec> open, pending;
    std::vector<hipEvent_t> free_ev;
    double total_ms[kNumKernelClasses] = {};
    int64_t count[kNumKernelClasses] = {};
    bool begin(int cls, hipStream_t st);   // false: this class is not timed
    void end(hipStream_t st);
    void collect();
    void reset();
    ~Profiler();
};

// Launch plan for one summation order (0 = pool, 1 = serial).
struct Plan {
    SysMeta *d_sys = nullptr;
    NPItem *d_items = nullptr;
    int2 *d_blk_pool = nullptr, *d_blk_serial = nullptr, *d_blk_recv = nullptr;
    int n_pool = 0, n_serial = 0, n_recv = 0;
    int nrows = 0, maxT = 1, max_n = 1;
    int64_t an_rows = 0;
    double *d_an = nullptr, *d_part = nullptr;
    int2 *d_work_pool = nullptr;     // (system, body) of every body of a pool-order system
    int n_work_pool = 0;
    bool np_fusable = true;          // every pool system fits the fused non-point pass
    int2 *d_blk16 = nullptr;         // (system, first body) of 16-body blocks, largest systems first
    int n_blk16 = 0;
    // item-field rows of the single-thread-order systems (see View::san)
    NPItem *d_sitems = nullptr;
    int n_sitems = 0;
    int32_t *d_srow_gp = nullptr, *d_srow_ring = nullptr;
    double *d_san = nullptr;
};

struct Batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n_sys = 0, nb = 0, nbp = 0;
    int max_sys_n = 0;
    std::vector<int64_t> sys_n, sys_off;
    std::vector<eph_geopotential> gps;
    std::vector<eph_ring> rings;
    std::vector<dd> t_eph;
    std::vector<uint64_t> sid;
    std::vector<int32_t> host_gp, host_ring;
    std::vector<double> host_k2, host_td;       // tidal class of each body
    std::vector<int64_t> host_gp_global, host_ring_global;

    double *F = nullptr;
    eph_geopotential *d_gps = nullptr;
    eph_ring *d_rings = nullptr;
    int32_t *d_np_gp = nullptr, *d_np_ring = nullptr, *d_bsys = nullptr;
    double *d_t0 = nullptr;
    int2 *d_blk_all = nullptr;
    int n_blk_all = 0;
    // deform work split by tidal class (see build_class_tables)
    int32_t *d_perm = nullptr;
    int4 *d_blk_light = nullptr, *d_blk_damped = nullptr;
    int n_blk_light = 0, n_blk_damped = 0;
    int2 *d_work_damped = nullptr;
    int n_work_damped = 0;
    int4 *d_blk_light16 = nullptr;
    int n_blk_light16 = 0;
    std::vector<void *> class_allocs;
    Plan plan[2];
    std::vector<void *> dev_allocs;
    Profiler profiler;
    bool force_generic = false;  // never take the persistent small path
    // kernel variant of the j-loop passes: 0 = automatic, 1 = one thread per
    // body, 2 = one workgroup per body (3: the same without the fused
    // non-point sums), 4 = four lanes per body
    int variant = 0;
    // wavefronts per body of the one-workgroup-per-body kernels: 0 =
    // automatic, else 1, 2 or 4
    int waves = 0;
    // automatic choices made by the autotuner (0: static heuristic)
    int tuned = 0, tuned_waves = 0;
    bool use_autotune = true, tune_done = false;
    int eff_variant() const { return variant ? variant : tuned; }
    int eff_waves() const { return waves ? waves : (variant ? 0 : tuned_waves); }
    int autotune(double dt, int mode);
    int2 *d_work_all = nullptr;
    int n_work_all = 0;
    bool use_wpb(int mode) const;

    ~Batch();
    int build_plan(int mode);
    int build_class_tables(const eph_mass *masses);
    View view(int mode) const;
    int accel(int mode, bool deform_only);
    int integrate(double dt, int64_t n_step, int mode);
    // persistent path: all systems serial and N <= kSmallMax; optional record
    int integrate_small(double dt, int64_t n_step, JMain *rec, int64_t rec_steps,
                        void (*after_chunk)(void *, int64_t, int64_t), void *ctx);
    bool all_small() const;
    std::vector<int> small_ids;
    int *d_small_ids = nullptr;
    int upload(const eph_mass *masses);
    int download(eph_mass *masses);
    // step graph (see step_graph)
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    double g_dt = 0;
    int g_mode = -1, g_variant = -1, g_waves = -1;
    int64_t *d_istep = nullptr;
    int step_graph(double dt, int mode);
    void drop_graph();
};

// Massless test particles advanced in the field of one small resident system.
struct Particles {
    Batch *field = nullptr;      // not owned; must outlive the particles
    int field_sys = 0;           // system index of the field in its batch
    int64_t np = 0, npp = 0;
    double GM = 0, R = 0;
    double *P = nullptr;
    JMain *rec = nullptr;        // two buffers of field records, rec_steps steps each
    int64_t rec_steps = 0;       // steps recorded per chunk
    int *d_sys_ids = nullptr;
    // particle launches run on their own stream (default priority), one
    // chunk behind the field's launches on the field's stream
    hipStream_t pstream = nullptr;
    hipEvent_t ev_rec[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
    ~Particles();
};

DtK stage_dtk();
int regroup_batch(Batch *old, const std::vector<int> &keep, int64_t n_new, const int64_t *new_sys_n,
                  const eph_mass *new_masses, const eph_geopotential *new_gps, int64_t n_new_gp,
                  const eph_ring *new_rings, int64_t n_new_ring, const double *new_t_eph, Batch **out);
int field_epoch_record(Batch *field, JMain *dst);
int create_particles(Batch *field, int64_t n, const double *r, const double *v, double GM, double R, Particles **out);
int integrate_particles(Particles *ps, double dt, int64_t n_step);
int download_particles(Particles *ps, double *r, double *v, double *extrema);

// eph_mass member of each field of the field-major layout, fields
// [0, kMassFieldEnd) except phi / c^2 and naccel / c^2
struct FieldMap { int field; size_t offset; };
constexpr int kMassFieldEnd = F_X0R;   // fields [0, F_X0R) carry eph_mass data
const std::vector<FieldMap> &mass_field_map();
// set-up of the allocation-free entry point on the current device
int noalloc_init_device(int device);

int create_batch(int64_t n_sys, const int64_t *sys_n, const eph_mass *masses,
                 const eph_geopotential *gps, int64_t n_gp, const eph_ring *rings, int64_t n_ring,
                 const double *t_eph, Batch **out);

}  // namespace eph
