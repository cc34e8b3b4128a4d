/*
 * ORACLE / golden-vector harness -- test infrastructure only, built into
 * oracle/_ref/ by `make ref` when the reference checkout is present.
 *
 * It compiles the reference's own translation unit src/physics/mass.cpp
 * (the thread-pool task bodies accel_deform, accel_mainforce, accel_nonpoint
 * and mass::deform_by, mass.cpp:78-214), links the reference's
 * geopotential.cpp and ring.cpp, and uses the force macros of
 * src/physics/mass.impl and the math headers unchanged.  Functions that would
 * need code this image cannot build (msystem::accel and RungeKutta12 call the
 * CUDA back end and the Windows-only recorder) are never referenced, and the
 * linker drops them.  The two drivers below reproduce the orchestration of
 * msystem::accel (pooled branch mass.cpp:226-307, single-thread branch
 * :309-374, finalisation :389-396) around those reference bodies.
 */
#include "physics/mass.cpp"
#include "eph_oracle.h"

#include <vector>

#define EXPORT extern "C" __attribute__((visibility("default")))

namespace {

real D(const o_dd &a) { return real(a.hi, a.lo); }
o_dd D(const real &a) { o_dd r = {a.hi, a.lo}; return r; }
mpvec DV(const o_ddv3 &a) { return mpvec(D(a.x), D(a.y), D(a.z)); }
o_ddv3 DV(const mpvec &a) { o_ddv3 r = {D(a.x), D(a.y), D(a.z)}; return r; }
vec V(const o_v3 &a) { return vec(a.x, a.y, a.z); }
o_v3 V(const vec &a) { o_v3 r = {a.x, a.y, a.z}; return r; }
mat M(const o_m3 &a) { return mat(V(a.x), V(a.y), V(a.z)); }
o_m3 M(const mat &a) { o_m3 r = {V(a.x), V(a.y), V(a.z)}; return r; }
mpmat DM(const o_ddm3 &a) { return mpmat(DV(a.x), DV(a.y), DV(a.z)); }
o_ddm3 DM(const mpmat &a) { o_ddm3 r = {DV(a.x), DV(a.y), DV(a.z)}; return r; }

struct models {
    std::vector<const geopotential *> gp;
    std::vector<const ring *> rg;
    models(const o_gp *gps, int64_t ngp, const o_ring *rings, int64_t nring) {
        for (int64_t k = 0; k < ngp; ++k) {
            /* same bytes geopotential::load builds: {Nz, Nt, c_table[]} */
            void *p = malloc(sizeof(o_gp));
            memcpy(p, &gps[k], sizeof(o_gp));
            gp.push_back((const geopotential *)p);
        }
        for (int64_t k = 0; k < nring; ++k) {
            const o_ring &o = rings[k];
            ring *r = (ring *)malloc(sizeof(ring) + 3 * sizeof(fast_real) * o.N);
            r->N = o.N; r->GM_ratio = o.GM_ratio; r->A = o.A; r->J2 = o.J2; r->GL_R2 = o.GL_R2;
            for (int64_t i = 0; i < o.N; ++i) {
                r->c_table[i].GsR2 = o.GsR2[i];
                r->c_table[i].R_R = o.R_R[i];
                r->c_table[i].H_R = o.H_R[i];
            }
            rg.push_back(r);
        }
    }
    ~models() {
        for (auto *p : gp) free((void *)p);
        for (auto *p : rg) free((void *)p);
    }
};

void to_mass(const o_mass &o, mass &m, const models &md) {
    m.r = DV(o.r); m.v = DV(o.v); m.GL = DV(o.GL); m.s = DM(o.s); m.w = V(o.w);
    m.GM = o.GM; m.exJ2 = o.exJ2; m.sid = o.sid;
    m.GM0 = o.GM0; m.dGM = o.dGM; m.dJ2 = o.dJ2; m.R = o.R; m.inertia = o.inertia;
    m.k2 = o.k2; m.k2r = o.k2r; m.tide_delay = o.tide_delay; m.tide_delay_factor = o.tide_delay_factor;
    m.lum = o.lum; m.recpt = o.recpt; m.A = o.A; m.R2 = o.R2; m.sR2 = o.sR2; m.rR2G_4c = o.rR2G_4c;
    m.C_static = M(o.C_static);
    m.gpmodel = o.gp >= 0 ? md.gp[o.gp] : nullptr;
    m.ringmodel = o.ring >= 0 ? md.rg[o.ring] : nullptr;
    m.phi = o.phi; m.beta2 = o.beta2;
    m.naccel = V(o.naccel); m.daccel = V(o.daccel); m.dtorque = V(o.dtorque);
    m.gaccel = V(o.gaccel); m.beta = V(o.beta);
    m.GI = M(o.GI); m.C_potential = M(o.C_potential);
    m.Egrad = V(o.Egrad); m.idaccel = V(o.idaccel); m.idtorque = V(o.idtorque);
    m.jdaccel = V(o.jdaccel); m.jdtorque = V(o.jdtorque);
    m.min_distance = o.min_distance; m.max_influence = o.max_influence;
}
void from_mass(const mass &m, o_mass &o) {
    o.r = DV(m.r); o.v = DV(m.v); o.GL = DV(m.GL); o.s = DM(m.s); o.w = V(m.w);
    o.GM = m.GM; o.exJ2 = m.exJ2;
    o.phi = m.phi; o.beta2 = m.beta2;
    o.naccel = V(m.naccel); o.daccel = V(m.daccel); o.dtorque = V(m.dtorque);
    o.gaccel = V(m.gaccel); o.beta = V(m.beta);
    o.GI = M(m.GI); o.C_potential = M(m.C_potential);
    o.Egrad = V(m.Egrad); o.idaccel = V(m.idaccel); o.idtorque = V(m.idtorque);
    o.jdaccel = V(m.jdaccel); o.jdtorque = V(m.jdtorque);
    o.min_distance = m.min_distance; o.max_influence = m.max_influence;
}

void finalize(htl::vector<mass> &mlist) {
    using Constants::c;
    using Constants::c2;
    int_t mn = mlist.size();
    for (int_t i = 0; i < mn; ++i) {
        mass &mi = mlist[i];
        FINALIZE_RELATIVITY;
        if (mi.ringmodel) {
            RING_CORRECTION;
        }
    }
}

/* pooled branch: the reference task bodies, called one task after another
 * (the pool runs them concurrently, but their write sets are disjoint) */
void accel_pooled(htl::vector<mass> &mlist) {
    int_t mn = mlist.size();
    const int_t n_tasklimit = 64, n_tasksize = 48 * 48;
    int_t n_task = std::min(mn, (mn * mn + (n_tasksize - 1)) / n_tasksize);
    n_task = std::min(n_task, n_tasklimit);
    for (int pass = 0; pass < 2; ++pass) {
        for (int_t t = 0; t < n_task; ++t) {
            int_t i_start = t * mn / n_task, i_end = (t + 1) * mn / n_task;
            mass &mi = mlist[i_start];
            mi.pmlist = &mlist;
            mi.task_index = i_end;
            if (pass == 0) accel_deform(&mi, 0);
            else accel_mainforce(&mi, 0);
        }
    }
    /* non-point items in creation order (last body first, geopotential item
     * before ring item); chain ch holds items [ch*n/T, (ch+1)*n/T) */
    struct item { int_t i; bool gp; };
    std::vector<item> items;
    for (int_t i = mn; i > 0;) {
        --i;
        mass &mi = mlist[i];
        if (mi.gpmodel) { items.push_back({i, true}); mi.idaccel = mi.idtorque = 0; }
        if (mi.ringmodel) { items.push_back({i, false}); mi.jdaccel = mi.jdtorque = 0; }
    }
    int_t n_np = items.size();
    int_t T = std::min(n_np, n_task);
    void *const end_marker = (char *)&mlist + 1;
    for (int_t k = 0; k < T; ++k) {
        for (int_t ch = 0; ch < T; ++ch) {
            int_t first = ch * n_np / T, last = (ch + 1) * n_np / T;
            void *prev = end_marker;
            for (int_t q = first; q < last; ++q) {
                mass &m = mlist[items[q].i];
                m.task_count = T;
                void **slot = items[q].gp ? &m.pmlist : &m.qmlist;
                *slot = prev;
                prev = slot;
            }
            mass &head = mlist[items[last - 1].i];
            if (items[last - 1].gp) head.task_index = ch + k;
            else head.task_jndex = ch + k;
            accel_nonpoint(prev, 0);
        }
    }
    if (n_np) for (auto &mi : mlist) {
        if (mi.gpmodel) { mi.daccel += mi.idaccel; mi.dtorque += mi.idtorque; }
        if (mi.ringmodel) { mi.daccel += mi.jdaccel; mi.dtorque += mi.jdtorque; }
    }
}

/* single-thread branch with the reference macros */
void accel_serial(htl::vector<mass> &mlist) {
    using Constants::c;
    using Constants::c2;
    int_t mn = mlist.size();
    for (int_t i = 0; i < mn; ++i) {
        mass &mi = mlist[i];
        int_t max_iter = MAX_ANGULAR_VELOCITY_ITER;
        do {
            mi.deform_by(mlist);
            bool should_break;
            UPDATE_ANGULAR_VELOCITY;
            if (should_break) break;
        } while (--max_iter);
        PREPARE_RELATIVITY;
        mi.daccel = mi.dtorque = mi.gaccel = 0;
    }
    for (int_t i = 0; i < mn; ++i) {
        mass &mi = mlist[i];
        for (int_t j = 0; j < mn; ++j) if (i != j) {
            mass &mj = mlist[j];
            RELATIVITY(mi);
            checked_maximize(mi.min_distance, rr);
            checked_maximize(mi.max_influence, tp_dg);
            ROTATIONAL_TIDAL_DEFORMATION;
            LENSE_THIRRING(mi);
            RADIATION_PRESSURE(mi);
            if (mi.gpmodel) {
                fast_mpmat fmis(mi.s);
                fast_mpvec lr = fmis.tolocal(r);
                fast_mpvec an = fmis.toworld(mi.gpmodel->sum(mi.R, lr));
                APPLY_NONPOINT_FORCE(mi);
            }
            if (mi.ringmodel) {
                fast_mpvec migl = mi.GL;
                fast_mpmat fgls(migl.asc_node(), 0, migl.unit());
                fast_mpvec lr = fgls.tolocal(r);
                fast_mpvec an = fgls.toworld(mi.ringmodel->sum(mi.R, lr));
                APPLY_NONPOINT_FORCE(mi);
            }
        }
    }
}

}  // namespace

/* Run one accel on n bodies: mode 0 = pooled branch (requires n_task > 1),
 * 1 = single-thread branch.  Bodies are updated in place. */
EXPORT int ref_accel(o_mass *ms, int64_t n, const o_gp *gps, int64_t ngp,
                     const o_ring *rings, int64_t nring, int mode) {
    models md(gps, ngp, rings, nring);
    htl::vector<mass> mlist;
    mlist.resize(n);
    for (int64_t i = 0; i < n; ++i) to_mass(ms[i], mlist[i], md);
    int_t n_task = std::min<int_t>(n, (n * n + 2303) / 2304);
    if (mode == 0) {
        if (n_task <= 1) return -1;
        accel_pooled(mlist);
    } else {
        accel_serial(mlist);
    }
    finalize(mlist);
    for (int64_t i = 0; i < n; ++i) from_mass(mlist[i], ms[i]);
    return 0;
}

EXPORT void ref_deform_by(o_mass *ms, int64_t n, int64_t i, const o_gp *gps, int64_t ngp,
                          const o_ring *rings, int64_t nring) {
    models md(gps, ngp, rings, nring);
    htl::vector<mass> mlist;
    mlist.resize(n);
    for (int64_t q = 0; q < n; ++q) to_mass(ms[q], mlist[q], md);
    mlist[i].deform_by(mlist);
    from_mass(mlist[i], ms[i]);
}

EXPORT void ref_gp_sum(const o_gp *gp, double R, const double *r, double *out) {
    models md(gp, 1, nullptr, 0);
    fast_mpvec g = md.gp[0]->sum(R, fast_mpvec(r[0], r[1], r[2]));
    out[0] = g.x; out[1] = g.y; out[2] = g.z;
}

EXPORT void ref_ring_sum(const o_ring *rg, double host_R, const double *r, double *out) {
    models md(nullptr, 0, rg, 1);
    fast_mpvec g = md.rg[0]->sum(host_R, fast_mpvec(r[0], r[1], r[2]));
    out[0] = g.x; out[1] = g.y; out[2] = g.z;
}

EXPORT void ref_dd_ops(const double *a, const double *b, double *out) {
    real x(a[0], a[1]), y(b[0], b[1]);
    real r[4] = {x + y, x - y, x * y, x / y};
    for (int k = 0; k < 4; ++k) { out[2 * k] = r[k].hi; out[2 * k + 1] = r[k].lo; }
}
EXPORT void ref_dd_sqrt(const double *a, double *out) {
    real r = sqrt(real(a[0], a[1]));
    out[0] = r.hi; out[1] = r.lo;
}
EXPORT void ref_rotation_matrix(const double *w, double t, double *out) {
    mat m = vec(w[0], w[1], w[2]).rotation_matrix(t);
    o_m3 o = M(m);
    memcpy(out, &o, sizeof o);
}
EXPORT void ref_inverse(const double *m9, double *out) {
    o_m3 i;
    memcpy(&i, m9, sizeof i);
    o_m3 o = M(M(i).inverse());
    memcpy(out, &o, sizeof o);
}
EXPORT double ref_normalize(double *a) {
    vec v(a[0], a[1], a[2]);
    double n = v.normalize();
    a[0] = v.x; a[1] = v.y; a[2] = v.z;
    return n;
}
EXPORT void ref_asc_node(const double *a, double *out) {
    vec v = vec(a[0], a[1], a[2]).asc_node();
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
