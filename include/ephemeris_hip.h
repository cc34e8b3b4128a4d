/*
 * ephemeris_hip.h -- C ABI of the MI355X (gfx950) force-evaluation and RK12
 * back end for Ephemeris.
 *
 * The library keeps systems resident in HBM.  One eph_system holds one or
 * more independent gravitating systems (a batch); every call computes all of
 * them.  All entry points return EPH_OK (0) or a negative status and set a
 * thread-local message readable with eph_last_error().  Nothing calls exit().
 * Calls on one eph_system must not overlap; different eph_system objects may
 * be used from different threads.
 *
 * Arithmetic reproduces the reference CPU path (double-double state, FP64
 * pair terms, no contraction) in the reference's summation order, so results
 * agree with it bit for bit apart from the last-ulp behaviour of libm
 * transcendentals (sin in the spin update; atan2/log/sin/cos in ring forces).
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   eph_accel       <- msystem::Cuda_accel()        src/integrators/RungeKutta.cu:418-438
 *                      (GPU twin of msystem::accel(int), src/physics/mass.cpp:216-397)
 *   eph_integrate   <- msystem::Cuda_RungeKutta12() src/integrators/RungeKutta.cu:391-412
 *                      msystem::integrate(dt,n,1)   src/physics/mass.cpp:399-442
 *   eph_deform      <- msystem::deform()            src/physics/mass.cpp:104-110
 *   eph_system_*    <- the AoS upload/download of cuda_rungekutta_kernel_config::load/save,
 *                      src/integrators/RungeKutta.cu:77-133 (here the state stays resident)
 */
#ifndef EPHEMERIS_HIP_H
#define EPHEMERIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPH_ABI_VERSION 1

enum {
    EPH_OK = 0,
    EPH_EINVAL = -1,      /* bad argument */
    EPH_ENODEV = -2,      /* no usable gfx950 device */
    EPH_ENOMEM = -3,      /* device or host allocation failed */
    EPH_EDEVICE = -4,     /* HIP runtime error */
    EPH_EIO = -5,         /* file not found / parse error */
    EPH_ESANITY = -6,     /* a body failed the loader sanity checks */
    EPH_ENOTSUP = -7,     /* requested back end not provided */
    EPH_EBUSY = -8        /* server: not enough free memory for a job now */
};

/* Summation order of one force evaluation (msystem::accel, mass.cpp:226-230):
 * POOL follows the thread-pool branch (taken when a pool exists and
 * min(N, ceil(N^2/2304)) > 1), SERIAL the single-thread branch. */
enum { EPH_SUMMATION_POOL = 0, EPH_SUMMATION_SERIAL = -1 };

/* One body: field-for-field mirror of class mass (src/physics/mass.h:67-143)
 * with double-double values as {hi, lo} and the model pointers replaced by
 * indices into the model arrays passed at creation (-1 = none).  Every field
 * is 8 bytes wide. */
typedef struct { double hi, lo; } eph_dd;
typedef struct { double x, y, z; } eph_vec;
typedef struct { eph_dd x, y, z; } eph_ddvec;
typedef struct { eph_vec x, y, z; } eph_mat;        /* three axes */
typedef struct { eph_ddvec x, y, z; } eph_ddmat;

typedef struct {
    eph_ddvec r, v, GL;                 /* position, velocity, G*L/GM */
    eph_ddmat s;                        /* body frame */
    eph_vec w;                          /* angular velocity */
    double GM, exJ2;                    /* time-variable parameters */
    uint64_t sid;                       /* id, up to 7 chars */
    double GM0, dGM, dJ2, R, inertia, k2, k2r, tide_delay, tide_delay_factor;
    double lum, recpt;
    double A, R2, sR2, rR2G_4c;         /* derived constants */
    eph_mat C_static;
    int64_t gp, ring;                   /* model indices, -1 = none */
    double phi, beta2;                  /* outputs of accel from here on */
    eph_vec naccel, daccel, dtorque, gaccel, beta;
    eph_mat GI, C_potential;
    eph_vec Egrad, idaccel, idtorque, jdaccel, jdtorque;
    double min_distance, max_influence;
} eph_mass;

/* Zonal/tesseral geopotential of degree 3..8 (src/physics/geopotential.h:4-38):
 * c[n*n-4] = J_n, c[n*n-4+2m-1] = C_nm, c[n*n-4+2m] = S_nm. */
#define EPH_GP_MAX_DEGREE 8
#define EPH_GP_TABLE 77
typedef struct { int64_t Nz, Nt; double c[EPH_GP_TABLE]; } eph_geopotential;

/* Ring of disks (src/physics/ring.h:4-28). */
#define EPH_RING_MAX_DISKS 64
typedef struct {
    int64_t N;
    double GM_ratio, A, J2, GL_R2;
    double GsR2[EPH_RING_MAX_DISKS], R_R[EPH_RING_MAX_DISKS], H_R[EPH_RING_MAX_DISKS];
} eph_ring;

typedef struct eph_system eph_system;

/* ---- library ---------------------------------------------------------- */
int eph_init(int device);                     /* select and check a gfx950 device */
const char *eph_last_error(void);
int eph_abi_version(void);

/* ---- systems resident in HBM ---------------------------------------------
 * A batch of n_sys systems: system k has sys_n[k] bodies, stored
 * consecutively in `masses`; model indices are global into gps/rings. */
int eph_system_create(int64_t n_sys, const int64_t *sys_n, const eph_mass *masses,
                      const eph_geopotential *gps, int64_t n_gp,
                      const eph_ring *rings, int64_t n_ring,
                      const double *t_eph /* n_sys x {hi,lo} */, eph_system **out);
int eph_system_destroy(eph_system *sys);
int eph_system_upload(eph_system *sys, const eph_mass *masses, const double *t_eph);
int eph_system_download(eph_system *sys, eph_mass *masses, double *t_eph);
int64_t eph_system_bodies(const eph_system *sys);
int64_t eph_system_count(const eph_system *sys);

/* ---- hot path ------------------------------------------------------------ */
int eph_deform(eph_system *sys);
int eph_accel(eph_system *sys, int summation);
/* msystem::integrate(dt, n_step, USE_GPU=1) for every system of the batch:
 * resets the running extrema, runs n_step RK12 steps with POOL summation,
 * inverts min_distance and advances t_eph by dt*n_step. */
int eph_integrate(eph_system *sys, double dt, int64_t n_step);
/* As eph_integrate but with an explicit summation order for every accel. */
int eph_integrate_ex(eph_system *sys, double dt, int64_t n_step, int summation);

/* Allocation-free force evaluation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc; it synchronises with the device, so it is
 * not a call to make under stream capture).  One msystem::accel
 * (parallel_option 0 with summation EPH_SUMMATION_POOL, -1 with
 * EPH_SUMMATION_SERIAL) of the n bodies at `masses` with their models, all in
 * device memory, written to `out` (device memory, n bodies; may be `masses`);
 * results are bit-identical to eph_system_create + eph_accel +
 * eph_system_download.  It calls no hipMalloc or hipFree: the working state
 * lives in a fixed array of the library on each device (about 2.8 MB, set up
 * by eph_init or by the first call on a device), shared by every call on that
 * device, so calls that overlap on one device race.  Limits: n <= 1024
 * bodies and at most 64 geopotential and ring models in use, else EPH_EINVAL.
 * The call runs on the null stream and returns when the results are in `out`;
 * inputs must be ready. */
int eph_accel_noalloc(const eph_mass *masses, int64_t n, const eph_geopotential *gps, int64_t n_gp,
                      const eph_ring *rings, int64_t n_ring, int summation, eph_mass *out);

/* ---- tuning ------------------------------------------------------------------
 * Kernel variant of the partner-loop passes: 0 automatic, 1 one thread per
 * body (throughput; large batches), 2 one wavefront per body (latency; few
 * bodies; the non-point pass runs inside the body's wavefront when every
 * thread-pool-order system has <= 1024 bodies and <= 64 non-point items),
 * 3 as 2 with the non-point pass in its own kernels, 4 four lanes per body
 * (throughput; the automatic choice for large batches).
 * allow_persistent = 0 keeps systems of <= 64 bodies off the
 * one-workgroup-per-system path.  All choices give bit-identical results.
 * No counterpart in the reference (its only knob is CUDA_CORES,
 * src/integrators/RungeKutta.cu:13-18, which changes results). */
int eph_system_set_variant(eph_system *sys, int variant, int allow_persistent);
/* Wavefronts per body of variants 2 and 3 (the j-tile of a round is 64 x
 * that many partners): 0 automatic (from the batch's shape), 1, 2 or 4.
 * The kernel with the fused non-point pass uses at least 2.  Results do not
 * depend on it.  eph_system_kernel_waves returns the value in use (0: chosen
 * per launch from the batch's shape).  No counterpart in the reference. */
int eph_system_set_waves(eph_system *sys, int waves_per_body);
int eph_system_kernel_waves(const eph_system *sys);
/* Autotuner (default on): in automatic mode, the first eph_integrate of a
 * batch with >= 2048 thread-pool-order bodies times the block shapes -- one
 * thread per body (variant 1), one workgroup of 1, 2 or 4 wavefronts per
 * body (variant 2; 2 and 4 only up to 8192 and 4096 bodies), four lanes per
 * body (variant 4) -- for a few steps on a device copy of the batch and
 * keeps the fastest (the choice is cached per batch shape for the process).
 * Results do not depend on it.  eph_system_kernel_variant returns the
 * variant in use (0: heuristic). */
int eph_system_set_autotune(eph_system *sys, int enable);
int eph_system_kernel_variant(const eph_system *sys);
/* enable = 1 (default): eph_integrate runs each RK12 step as one replay of a
 * captured HIP graph (25 stage updates and force evaluations) whenever the
 * profiler is off and n_step >= 2; 0 launches every kernel from the host.
 * Results are bit-identical either way. */
int eph_system_set_graph(eph_system *sys, int enable);

/* ---- massless test particles ------------------------------------------------
 * n particles, each integrated as one more body (appended after the field
 * bodies) of the `field` system -- which must be a batch of one system of at
 * most 48 bodies, so the reference's single-thread summation order applies.
 * A particle exerts no force (its terms are below the field bodies' last
 * ulp); GM and R are its gravitational parameter and radius.  r and v are
 * n x 3 FP64 (double-double with lo = 0).  eph_particles_integrate advances
 * the field system and the particles together; the field must outlive the
 * particle set.  Downloads: r, v as n x 6 {x.hi,x.lo,y.hi,...}; extrema as
 * n x {min_distance (running max of 1/r), max_influence}.
 * No counterpart in the reference's interface: the reference would carry
 * each particle as an ordinary body (mass_loader.cpp:324-471). */
typedef struct eph_particles eph_particles;
int eph_particles_create(eph_system *field, int64_t n, const double *r, const double *v, double GM, double R,
                         eph_particles **out);
int eph_particles_integrate(eph_particles *ps, double dt, int64_t n_step);
int eph_particles_download(eph_particles *ps, double *r, double *v, double *extrema);
int64_t eph_particles_count(const eph_particles *ps);
int eph_particles_destroy(eph_particles *ps);

/* ---- serving -----------------------------------------------------------------
 * A request loop over resident state: jobs are submitted at any time (from
 * any thread), admitted first-in first-out at tick boundaries while their
 * estimated HBM footprint fits the budget and the device's free memory, and
 * advanced together -- every admitted ensemble system in one batch, every
 * admitted particle set in one launch over a shared pool -- until their span
 * is done; state moves between batches device to device.
 *  - Ensemble job: n_sys independent systems (model indices local to the
 *    job), n_steps RK12 steps of cfg.ens_dt each, taken in calls of
 *    cfg.ens_quantum steps (n_steps must be a multiple of it): the result is
 *    bitwise that of eph_integrate(sys, ens_dt, ens_quantum) repeated.
 *  - Particle job: n particles of radius cfg.part_R in the server's field
 *    system from its epoch, n_steps steps of cfg.part_dt (at most
 *    cfg.part_max_steps): bitwise that of eph_particles_create on a copy of
 *    the field followed by eph_particles_integrate(ps, part_dt, n_steps).
 * Jobs that fail validation at submission are rejected (EPH_EINVAL /
 * EPH_ESANITY, no job id); a job whose state turns non-finite ends FAILED
 * without touching other jobs, whose results never depend on what else is
 * resident.  Times are seconds on a steady clock since eph_server_create.
 * eph_server_tick runs one tick on the calling thread; eph_server_start runs
 * ticks on a worker thread until eph_server_stop (eph_server_tick then
 * returns EPH_EBUSY).
 * No counterpart in the reference, which integrates one system per process
 * (src/modules/ephemeris_generator.cpp:130-166). */
typedef struct eph_server eph_server;
typedef struct {
    int64_t mem_budget;      /* bytes for resident jobs; 0 = 40% of free device memory */
    double ens_dt;           /* ensemble step (s) */
    int64_t ens_quantum;     /* ensemble steps per tick */
    double part_dt;          /* particle step (s) */
    int64_t part_quantum;    /* particle steps per tick */
    int64_t part_max_steps;  /* longest particle span (field records are kept for it) */
    double part_GM, part_R;  /* particle gravitational parameter and radius */
    int64_t max_systems;     /* ensemble slots; 0 = unlimited */
    int64_t max_particles;   /* particle slots; 0 = unlimited */
} eph_server_config;
enum { EPH_JOB_QUEUED = 0, EPH_JOB_RUNNING = 1, EPH_JOB_DONE = 2, EPH_JOB_FAILED = 3 };
enum { EPH_JOB_ENSEMBLE = 0, EPH_JOB_PARTICLES = 1 };
typedef struct {
    int32_t kind, status;
    int64_t steps_total, steps_done, bytes, units;   /* units: bodies or particles */
    double t_submit, t_admit, t_done;                 /* -1 until reached */
    char error[160];
} eph_job_info;
typedef struct {
    int64_t ticks, jobs_submitted, jobs_done, jobs_failed, jobs_queued, jobs_running;
    int64_t mem_used, mem_peak, mem_budget;
    double pair_interactions, rk12_system_steps, particle_steps, busy_s, t_now;
} eph_server_stats;
/* field: one system of at most 48 bodies (copied; NULL = no particle jobs) */
int eph_server_create(const eph_server_config *cfg, eph_system *field, eph_server **out);
int eph_server_submit_ensemble(eph_server *srv, int64_t n_sys, const int64_t *sys_n, const eph_mass *masses,
                               const eph_geopotential *gps, int64_t n_gp, const eph_ring *rings,
                               int64_t n_ring, const double *t_eph, int64_t n_steps, int64_t *job);
int eph_server_submit_particles(eph_server *srv, int64_t n, const double *r, const double *v,
                                int64_t n_steps, int64_t *job);
int eph_server_tick(eph_server *srv, int64_t *n_active);
int eph_server_start(eph_server *srv);
int eph_server_wait(eph_server *srv, double timeout_s);     /* until no job is queued or running */
int eph_server_stop(eph_server *srv);
int eph_server_job_info(eph_server *srv, int64_t job, eph_job_info *info);
/* results of a DONE job, layouts as eph_system_download / eph_particles_download */
int eph_server_result_ensemble(eph_server *srv, int64_t job, eph_mass *masses, double *t_eph);
int eph_server_result_particles(eph_server *srv, int64_t job, double *r, double *v, double *extrema);
/* State of an ensemble job at the last whole quantum (queued: its input;
 * running: read from the resident batch between ticks, waiting at most for
 * the tick in progress; done: its result), with the steps done so far.  A caller writes it as the reference's binary
 * checkpoint (eph_msystem_create, eph_msystem_analyse with reconstruct = 1,
 * eph_msystem_save_checkpoint; one file per member) and resumes from it by
 * loading the file and submitting the rest of the span.  No counterpart in
 * the reference, whose checkpoints are written by msystem::save_checkpoint
 * (src/modules/mass_loader.cpp:715-759). */
int eph_server_snapshot_ensemble(eph_server *srv, int64_t job, eph_mass *masses, double *t_eph, int64_t *steps_done);
int eph_server_release(eph_server *srv, int64_t job);       /* forget a finished job */
int eph_server_stats_read(eph_server *srv, eph_server_stats *st);
int eph_server_destroy(eph_server *srv);

/* ---- measurement ---------------------------------------------------------- */
/* Average device time (ms) per launch of each kernel class, measured with
 * HIP events on the library's stream since the last reset; names[] gets
 * static strings.  Returns the number of classes written. */
int eph_profile_enable(eph_system *sys, int enable);
/* Time only the classes whose bit (1 << index in the names of
 * eph_profile_read) is set; all by default.  Each timed launch adds two
 * event records to the stream. */
int eph_profile_classes(eph_system *sys, unsigned mask);
int eph_profile_read(eph_system *sys, const char **names, double *avg_ms, int64_t *launches, int cap);
int eph_profile_reset(eph_system *sys);

/* ---- host system object (the msystem operator interface) ------------------
 * The msystem of the reference around the calls above; the host copy of the
 * bodies is authoritative, the GPU keeps a resident copy refreshed only when
 * the host copy changed.
 *   eph_msystem_load            <- msystem::load(fconfig)   src/modules/mass_loader.cpp:67-294, 324-561
 *                                  (TDB correction :225-237; initial deform, C_static adjustment,
 *                                  accel and analyse(true) :540-558, on the CPU back end)
 *   eph_msystem_create          <- msystem::push_back of each body + copy_params (mass_loader.cpp:924-1004)
 *   eph_msystem_load_checkpoint <- msystem::load_checkpoint  mass_loader.cpp:614-698
 *   eph_msystem_save_checkpoint <- msystem::save_checkpoint  mass_loader.cpp:715-759
 *                                  (binary layout of the reference: header, barycen ids, per body the
 *                                  824 bytes of class mass up to Egrad with model pointers replaced
 *                                  by the models' byte counts, then the geopotential and ring objects)
 *   eph_msystem_accel           <- msystem::accel(int)       src/physics/mass.cpp:216-397
 *                                  parallel_option 1 = GPU (msystem::Cuda_accel, RungeKutta.cu:418-438),
 *                                  0 = CPU threads (the pool order when min(N, ceil(N^2/2304)) > 1),
 *                                  -1 = CPU single-thread order; the CPU accel applies the tidal-matrix
 *                                  correction (mass.cpp:376-387), the GPU accel does not (as in the reference)
 *   eph_msystem_integrate       <- msystem::integrate(dt,n,USE_GPU) mass.cpp:399-442
 *                                  USE_GPU 0 = CPU RK12 (RungeKutta.impl:107-194), 1 = GPU RK12
 *   eph_msystem_combined_integrate <- msystem::combined_integrate src/integrators/mass_combined.cpp:40-406
 *   eph_msystem_update          <- msystem::update(t)        mass.cpp:28-48
 *   eph_msystem_analyse         <- msystem::analyse(reconstruct) src/physics/barycen_analyzer.cpp:12-310
 *   eph_msystem_scale / _scale_geopotential / _scale_ring
 *                               <- msystem::scale / scale_geopotential / scale_ring, mass_loader.cpp:1005-1026
 *   eph_msystem_set_tidal       <- the tidal_parent / tidal_matrix / tidal_childlist members (mass.h:236-240)
 *                                  that combined_integrate sets on its subsystems
 * Diagnostics mirror the collision (kind 0: value = the closest approach,
 * limit = R) and step-size (kind 1: value = |dt|, limit = recommended dt)
 * reports of mass.cpp:411-439; eph_msystem_diagnostics returns their number
 * and copies up to cap of them.  eph_msystem_barycens writes per barycen
 * {pid, hid, gid, tid, mid, number of children, offset of its children in
 * `children`} (7 int64), GM_sys as {hi, lo}, and the children lists, and
 * returns the number of barycens. */
typedef struct eph_msystem eph_msystem;
int eph_msystem_load(const char *config_path, int summation_at_load, eph_msystem **out);
/* as eph_msystem_load; analyse = 0 skips the loader's orbit-structure
 * analysis (for systems whose structure is never used: ensemble members,
 * particle fields).  No counterpart in the reference, whose loader always
 * analyses (mass_loader.cpp:558). */
int eph_msystem_load_ex(const char *config_path, int summation_at_load, int analyse, eph_msystem **out);
int eph_msystem_create(int64_t n, const eph_mass *masses, const eph_geopotential *gps, int64_t n_gp,
                       const eph_ring *rings, int64_t n_ring, const double *t_eph, eph_msystem **out);
int eph_msystem_load_checkpoint(const char *path, eph_msystem **out);
int eph_msystem_save_checkpoint(const eph_msystem *h, const char *path);
int eph_msystem_destroy(eph_msystem *h);
int64_t eph_msystem_size(const eph_msystem *h);
int eph_msystem_masses(const eph_msystem *h, eph_mass *out, int64_t cap);
int eph_msystem_set_masses(eph_msystem *h, const eph_mass *masses, int64_t n);
int eph_msystem_models(const eph_msystem *h, eph_geopotential *gps, int64_t gcap, eph_ring *rings,
                       int64_t rcap, int64_t *n_gp, int64_t *n_ring);
int eph_msystem_time(const eph_msystem *h, double *t_eph);
int eph_msystem_set_time(eph_msystem *h, const double *t_eph);
/* delta_t, data_cadence, max_ephm_length, combined_delta_t, GM_max_child,
 * GM_max_parent, GM_max_tiny, Period_max_child (mass.h:214-232) */
int eph_msystem_params(const eph_msystem *h, double *params8, int64_t *integrator);
int eph_msystem_accel(eph_msystem *h, int parallel_option);
int eph_msystem_integrate(eph_msystem *h, double dt, int64_t n_step, int use_gpu);
int eph_msystem_combined_integrate(eph_msystem *h, double dt, int64_t n_combine, int64_t n_step, int use_gpu);
int eph_msystem_update(eph_msystem *h, double t);
int eph_msystem_analyse(eph_msystem *h, int reconstruct, double *t_update);
int64_t eph_msystem_barycens(const eph_msystem *h, int64_t *ids, double *gm_sys, int64_t *children, int64_t cap,
                             int64_t child_cap);
int eph_msystem_scale(eph_msystem *h, int64_t mid, double factor);
int eph_msystem_scale_geopotential(eph_msystem *h, int64_t mid, double factor);
int eph_msystem_scale_ring(eph_msystem *h, int64_t mid, double factor);
int eph_msystem_set_tidal(eph_msystem *h, int64_t parent, const double *matrix9, const int64_t *children,
                          int64_t n_children);
/* worker threads of the CPU back end (0 = all hardware threads); results do
 * not depend on it */
int eph_msystem_set_threads(eph_msystem *h, int n_threads);
int64_t eph_msystem_get_mid(const eph_msystem *h, const char *ssid);
int eph_msystem_diagnostics(const eph_msystem *h, int64_t *body, int *kind, double *value,
                            double *limit, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
