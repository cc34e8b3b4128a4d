/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference CPU path of Ephemeris v0.4.2 for the
 * force evaluation (msystem::accel) and the 25-stage RK12 integrator, plus the
 * text-format loader that produces identical inputs.  It exists to check the
 * HIP implementation and to time a CPU baseline.  Only tests/, the smoke entry
 * point and the bench CPU-baseline leg may load it; the product never does.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * the reference checkout).  Arithmetic follows the reference expression by
 * expression with FP contraction off, so that a correct HIP kernel that keeps
 * the same summation order reproduces it bit for bit (libm transcendentals
 * excepted).
 *
 * Parity pinning of each part is listed in oracle/README.md.
 */
#ifndef EPH_ORACLE_H
#define EPH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double hi, lo; } o_dd;
typedef struct { double x, y, z; } o_v3;
typedef struct { o_dd x, y, z; } o_ddv3;
typedef struct { o_v3 x, y, z; } o_m3;      /* three axes (columns) */
typedef struct { o_ddv3 x, y, z; } o_ddm3;

/* Mirror of class mass (src/physics/mass.h:67-160); every field is 8 bytes so
 * a numpy structured dtype can view an array of them. */
typedef struct {
    o_ddv3 r, v, GL;
    o_ddm3 s;
    o_v3 w;
    double GM, exJ2;
    uint64_t sid;
    double GM0, dGM, dJ2, R, inertia, k2, k2r, tide_delay, tide_delay_factor;
    double lum, recpt;
    double A, R2, sR2, rR2G_4c;
    o_m3 C_static;
    int64_t gp, ring;               /* model index, -1 when absent */
    double phi, beta2;
    o_v3 naccel, daccel, dtorque, gaccel, beta;
    o_m3 GI, C_potential;
    o_v3 Egrad, idaccel, idtorque, jdaccel, jdtorque;
    double min_distance, max_influence;
} o_mass;

#define O_GP_MAXN 8
#define O_GP_TABLE (O_GP_MAXN * (O_GP_MAXN + 2) - 3)
typedef struct { int64_t Nz, Nt; double c[O_GP_TABLE]; } o_gp;

#define O_RING_MAXDISK 64
typedef struct {
    int64_t N;
    double GM_ratio, A, J2, GL_R2;
    double GsR2[O_RING_MAXDISK], R_R[O_RING_MAXDISK], H_R[O_RING_MAXDISK];
} o_ring;

typedef struct o_system {
    int64_t n;
    o_mass *m;
    o_dd t_eph;
    double delta_t;
    int64_t n_gp, n_ring;
    o_gp *gp;
    o_ring *ring;
    int nthreads;
    char err[256];
} o_system;

/* --- system life cycle ------------------------------------------------- */
o_system *o_load_config(const char *config_path, int pool_at_load);
o_system *o_load_files(const char *initial, const char *extra,
                       const char *gp_dir, const char *ring_dir, int pool_at_load);
o_system *o_clone(const o_system *s);
/* keep the subset idx[0..k) (in that order), as msystem::push_back does */
o_system *o_subset(const o_system *s, const int64_t *idx, int64_t k);
void o_free(o_system *s);
const char *o_load_error(void);
int64_t o_size(const o_system *s);
o_mass *o_masses(o_system *s);
int o_set_threads(o_system *s, int nthreads);

/* --- hot path ------------------------------------------------------------ */
/* msystem::accel(parallel_option) with a thread pool present or not
 * (src/physics/mass.cpp:216-397). */
void o_accel(o_system *s, int parallel_option, int pool);
/* msystem::deform (src/physics/mass.cpp:104-110) */
void o_deform(o_system *s);
/* msystem::integrate(dt,n_step,0) (src/physics/mass.cpp:399-442) */
void o_integrate(o_system *s, double dt, int64_t n_step, int pool);
/* msystem::update (src/physics/mass.cpp:28-48) */
void o_update(o_system *s, double t);
/* mass::scale (mass.cpp:8-26) and msystem::scale_geopotential / scale_ring
 * (mass_loader.cpp:1005-1026) */
void o_scale(o_system *s, int64_t mid, double factor);
void o_scale_geopotential(o_system *s, int64_t mid, double factor);
void o_scale_ring(o_system *s, int64_t mid, double factor);

/* --- single functions, exported for golden-vector checks -------------------- */
void o_dd_ops(const double *a, const double *b, double *out /* 8: +,-,*,/ */);
void o_dd_sqrt(const double *a, double *out);
void o_rotation_matrix(const double *w, double t, double *out9);
void o_inverse(const double *m9, double *out9);
double o_normalize(double *v3);
void o_asc_node(const double *v, double *out3);
void o_gp_sum(const o_gp *gp, double R, const double *r, double *out3);
void o_ring_sum(const o_ring *rg, double host_R, const double *r, double *out3);
float o_fzcorr(float x, float y, int kh);
/* deform_by on masses[i] against the array (src/physics/mass.cpp:78-103) */
void o_deform_by(o_mass *m, int64_t n, int64_t i);
/* RK12 tableau (flat lower triangle, hi/lo pairs) */
const double *o_rk12_table(void);

#ifdef __cplusplus
}
#endif
#endif
