// Host (CPU-thread) back end: see eph_cpu.hpp.  Host code only; the pair
// terms come from eph_terms.hpp, which the gfx950 kernels compile as well.
#include "eph_cpu.hpp"

#include <algorithm>
#include <cstring>

#include "eph_terms.hpp"

namespace eph {

// ---- worker pool --------------------------------------------------------------
WorkerPool::WorkerPool(int n_threads) {
    for (int t = 1; t < std::max(1, n_threads); ++t) workers.emplace_back([this] { work(); });
}

WorkerPool::~WorkerPool() {
    {
        std::lock_guard<std::mutex> lk(mu);
        quit = true;
    }
    cv_go.notify_all();
    for (auto &t : workers) t.join();
}

void WorkerPool::work() {
    uint64_t seen = 0;
    for (;;) {
        const std::function<void(int64_t)> *fn;
        uint64_t gen;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_go.wait(lk, [&] { return quit || generation != seen; });
            if (quit) return;
            seen = gen = generation;
            fn = job;
        }
        for (;;) {
            int64_t i;
            {
                // an index is taken only from the job this worker woke for:
                // run() of a later job cannot start before every index of
                // this one is done, and generation tells the jobs apart
                std::lock_guard<std::mutex> lk(mu);
                if (generation != gen || next >= job_n) break;
                i = next++;
            }
            (*fn)(i);
            std::lock_guard<std::mutex> lk(mu);
            if (++done == job_n) cv_done.notify_all();
        }
    }
}

void WorkerPool::run(int64_t n, const std::function<void(int64_t)> &fn) {
    if (n <= 0) return;
    if (workers.empty() || n == 1) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    uint64_t gen;
    {
        std::lock_guard<std::mutex> lk(mu);
        job = &fn;
        job_n = n;
        next = 0;
        done = 0;
        gen = ++generation;
    }
    cv_go.notify_all();
    for (;;) {                                          // the caller works too
        int64_t i;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (next >= job_n) break;
            i = next++;
        }
        fn(i);
        std::lock_guard<std::mutex> lk(mu);
        if (++done == job_n) cv_done.notify_all();
    }
    std::unique_lock<std::mutex> lk(mu);
    cv_done.wait(lk, [&] { return done == job_n; });
    (void)gen;
    job = nullptr;
}

namespace {

// ---- host copies of the ring model's constant tables (ring.impl:19-2068, 2100-2107)
const float h_fz[32 * 64 * 16] = {
#include "../tables/ring_fzcorr.inc"
};
#include "../tables/ring_pade.inc"
const double h_xserkk[][2] = EPH_PADE_XSERKK;
const double h_xserkm[][2] = EPH_PADE_XSERKM;
const double h_xser0[][2] = EPH_PADE_XSER0;
const double h_kserkk[][2] = EPH_PADE_KSERKK;
const double h_kserkm[][2] = EPH_PADE_KSERKM;
const double h_kser0[][2] = EPH_PADE_KSER0;

struct HostRingTab {
    static float fz(float x, float y, int kh) { return fzcorr_tab(h_fz, x, y, kh); }
    static double xserkk(double x) { return padesum(x, h_xserkk); }
    static double xserkm(double x) { return padesum(x, h_xserkm); }
    static double xser0(double x) { return padesum(x, h_xser0); }
    static double kserkk(double x) { return padesum(x, h_kserkk); }
    static double kserkm(double x) { return padesum(x, h_kserkm); }
    static double kser0(double x) { return padesum(x, h_kser0); }
};

const double kRKTable[325][2] = {
#include "../tables/rk12_feagin.inc"
};

struct HostSin { double operator()(double x) const { return std::sin(x); } };

v3 V(const eph_vec &a) { return v3{a.x, a.y, a.z}; }
eph_vec E(v3 a) { return eph_vec{a.x, a.y, a.z}; }
m3 M(const eph_mat &a) { return m3{V(a.x), V(a.y), V(a.z)}; }
eph_mat E(const m3 &a) { return eph_mat{E(a.x), E(a.y), E(a.z)}; }
ddv3 DV(const eph_ddvec &a) { return ddv3{dd{a.x.hi, a.x.lo}, dd{a.y.hi, a.y.lo}, dd{a.z.hi, a.z.lo}}; }
eph_ddvec E(const ddv3 &a) { return eph_ddvec{{a.x.hi, a.x.lo}, {a.y.hi, a.y.lo}, {a.z.hi, a.z.lo}}; }
ddm3 DM(const eph_ddmat &a) { return ddm3{DV(a.x), DV(a.y), DV(a.z)}; }
eph_ddmat E(const ddm3 &a) { return eph_ddmat{E(a.x), E(a.y), E(a.z)}; }

constexpr int kMaxSpinIter = 4;                      // MAX_ANGULAR_VELOCITY_ITER, mass.impl:74

void run(WorkerPool *pool, int64_t n, const std::function<void(int64_t)> &fn) {
    if (pool) pool->run(n, fn);
    else for (int64_t i = 0; i < n; ++i) fn(i);
}

// partner sum of deform_by for body i with spin w
DeformSum deform_sum(const CpuSystem &s, int64_t i, int cls, v3 w) {
    const eph_mass &mi = s.m[i];
    const ddv3 ri = DV(mi.r), vi = DV(mi.v);
    DeformSum sum{0, v3{0, 0, 0}, mzero()};
    for (int64_t j = 0; j < s.n; ++j) {
        if (j == i) continue;
        const eph_mass &mj = s.m[j];
        deform_pair(cls, ri, vi, mi.sR2, mi.k2, mi.tide_delay, w, DV(mj.r), DV(mj.v), mj.sR2, mj.GM,
                    mj.tide_delay_factor, sum);
    }
    return sum;
}

Harmonics harmonics_of(const eph_mass &mi) {
    Harmonics h;
    h.set(mi.k2r, mi.R, mi.R2, mi.A, mi.GM, mi.exJ2, M(mi.C_static), todbl(DM(mi.s)));
    return h;
}

// accel_deform's per-body work (mass.cpp:112-134, 311-326): deform passes
// with the spin fixed point, PREPARE_RELATIVITY, reset of the accumulators.
// The partner sum of a body without constant-time lag does not depend on the
// spin, so it is formed once for all passes (same values, same bits).
void deform_spin_prepare(const CpuSystem &s, int64_t i) {
    eph_mass &mi = s.m[i];
    const int cls = tidal_class(mi.k2, mi.tide_delay);
    const Harmonics H = harmonics_of(mi);
    v3 w = V(mi.w);
    ddv3 GL = DV(mi.GL);
    DeformSum sum = deform_sum(s, i, cls, w);
    m3 C, GI;
    for (int iter = kMaxSpinIter;;) {
        H.apply(sum.C, w, C, GI);
        bool gl_from_w;
        const bool brk = spin_step(GI, w, GL, gl_from_w);
        if (gl_from_w) GL = mkddv(GI * w);
        if (brk || --iter == 0) break;
        if (cls == 2) sum = deform_sum(s, i, cls, w);
    }
    mi.GL = E(GL);
    mi.w = E(w);
    mi.C_potential = E(C);
    mi.GI = E(GI);
    const v3 beta = todbl(DV(mi.v)) / kC;              // PREPARE_RELATIVITY, mass.impl:76-80
    mi.beta = E(beta);
    mi.beta2 = dot(beta, beta);
    mi.phi = sum.phi / kC2;
    mi.naccel = E(sum.nacc / kC2);
    mi.daccel = mi.dtorque = mi.gaccel = eph_vec{0, 0, 0};
}

IMain imain_of(const eph_mass &m) {
    IMain I;
    I.r = DV(m.r);
    I.sR2 = m.sR2; I.phi = m.phi; I.beta2 = m.beta2;
    I.R2 = m.R2; I.GM = m.GM; I.rR2G4c = m.rR2G_4c;
    I.beta = V(m.beta);
    I.GL = todbl(DV(m.GL));
    I.C = M(m.C_potential);
    return I;
}

JMain jmain_of(const eph_mass &m) {
    JMain T;
    memset(&T, 0, sizeof T);
    T.rxh = m.r.x.hi; T.rxl = m.r.x.lo; T.ryh = m.r.y.hi; T.ryl = m.r.y.lo; T.rzh = m.r.z.hi; T.rzl = m.r.z.lo;
    T.GM = m.GM; T.sR2 = m.sR2;
    T.phi = m.phi; T.beta2 = m.beta2;
    T.bx = m.beta.x; T.by = m.beta.y; T.bz = m.beta.z;
    T.nx = m.naccel.x; T.ny = m.naccel.y; T.nz = m.naccel.z;
    const double *C = &m.C_potential.x.x;
    for (int q = 0; q < 9; ++q) T.C[q] = C[q];
    T.R2 = m.R2;
    const v3 gl = todbl(DV(m.GL));
    T.GLx = gl.x; T.GLy = gl.y; T.GLz = gl.z;
    T.lum = m.lum;
    return T;
}

// field of body p's geopotential (kind 0) or ring (kind 1) at offset r, per unit GM
v3 nonpoint_field(const CpuSystem &s, const eph_mass &mp, int kind, v3 r) {
    if (kind == 0) return gp_field_world(s.gps[mp.gp], mp.R, todbl(DM(mp.s)), r);
    return ring_field_world<HostRingTab>(s.rings[mp.ring], mp.R, todbl(DV(mp.GL)), r);
}

// Non-point rounds of the thread-pool branch (mass.cpp:253-307, task body
// accel_nonpoint :158-214).  Items are created from the last body to the
// first, the geopotential before the ring of the same body; chain c holds
// items [c nnp / T, (c + 1) nnp / T) and walks them from its last-created
// item back to its first; in round k it covers partner slice (c + k) mod T.
// The chains of one round touch disjoint slices, so they run in parallel.
void nonpoint_rounds(const CpuSystem &s, int64_t n_task, WorkerPool *pool) {
    const int64_t mn = s.n;
    int64_t nnp = 0;
    for (int64_t i = 0; i < mn; ++i) {
        eph_mass &mi = s.m[i];
        if (mi.gp >= 0) { ++nnp; mi.idaccel = mi.idtorque = eph_vec{0, 0, 0}; }
        if (mi.ring >= 0) { ++nnp; mi.jdaccel = mi.jdtorque = eph_vec{0, 0, 0}; }
    }
    const int64_t T = std::min(nnp, n_task);
    if (T > 0) {
        struct Item { int64_t i; int kind; };
        std::vector<Item> items;
        for (int64_t i = mn; i > 0;) {
            --i;
            if (s.m[i].gp >= 0) items.push_back(Item{i, 0});
            if (s.m[i].ring >= 0) items.push_back(Item{i, 1});
        }
        for (int64_t k = 0; k < T; ++k) {
            run(pool, T, [&](int64_t c) {
                const int64_t first = c * nnp / T, last = (c + 1) * nnp / T;
                int64_t sl = c + k;
                if (sl >= T) sl -= T;
                const int64_t j0 = sl * mn / T, j1 = (sl + 1) * mn / T;
                for (int64_t q = last - 1; q >= first; --q) {
                    const int64_t i = items[q].i;
                    eph_mass &mi = s.m[i];
                    v3 td{0, 0, 0}, tq{0, 0, 0};
                    for (int64_t j = j0; j < j1; ++j) {
                        if (j == i) continue;
                        eph_mass &mj = s.m[j];
                        const v3 r = diff(DV(mj.r), DV(mi.r));
                        const v3 an = nonpoint_field(s, mi, items[q].kind, r);
                        mj.daccel = E(V(mj.daccel) + mi.GM * an);        // APPLY_NONPOINT_FORCE, mass.impl:127-130
                        td -= mj.GM * an;
                        tq -= mj.GM * cross(r, an);
                    }
                    if (items[q].kind == 0) {
                        mi.idaccel = E(V(mi.idaccel) + td);
                        mi.idtorque = E(V(mi.idtorque) + tq);
                    } else {
                        mi.jdaccel = E(V(mi.jdaccel) + td);
                        mi.jdtorque = E(V(mi.jdtorque) + tq);
                    }
                }
            });
        }
    }
    if (nnp)
        for (int64_t i = 0; i < mn; ++i) {
            eph_mass &mi = s.m[i];
            if (mi.gp >= 0) {
                mi.daccel = E(V(mi.daccel) + V(mi.idaccel));
                mi.dtorque = E(V(mi.dtorque) + V(mi.idtorque));
            }
            if (mi.ring >= 0) {
                mi.daccel = E(V(mi.daccel) + V(mi.jdaccel));
                mi.dtorque = E(V(mi.dtorque) + V(mi.jdtorque));
            }
        }
}

// single-thread branch of the main force with action-reaction updates
// (mass.cpp:328-373): partner j's daccel is updated in place
void main_serial(const CpuSystem &s, const std::vector<JMain> &J) {
    for (int64_t i = 0; i < s.n; ++i) {
        eph_mass &mi = s.m[i];
        const IMain I = imain_of(mi);
        v3 ga = V(mi.gaccel), da = V(mi.daccel), dq = V(mi.dtorque);
        double mind = mi.min_distance, maxi = mi.max_influence;
        for (int64_t j = 0; j < s.n; ++j) {
            if (j == i) continue;
            eph_mass &mj = s.m[j];
            const JMain &T = J[j];
            const PairGeo p = relativity(I, T, ga);            // RELATIVITY(mi)
            mind = checked_max_into(mind, p.rr);
            maxi = checked_max_into(maxi, p.tp_dg);
            const v3 dg = (p.rr3 * p.rr2 * I.R2) * quad_field(I.C, p);   // ROTATIONAL_TIDAL_DEFORMATION
            da += T.GM * dg;
            mj.daccel = E(V(mj.daccel) - mi.GM * dg);
            dq += T.GM * cross(p.r, dg);
            lt_rad(I, T, p, da, dq);                            // LENSE_THIRRING, RADIATION_PRESSURE
            for (int kind = 0; kind < 2; ++kind) {              // geopotential, then ring
                if ((kind == 0 ? mi.gp : mi.ring) < 0) continue;
                const v3 an = nonpoint_field(s, mi, kind, p.r);
                mj.daccel = E(V(mj.daccel) + mi.GM * an);       // APPLY_NONPOINT_FORCE(mi)
                da -= T.GM * an;
                dq -= T.GM * cross(p.r, an);
            }
        }
        mi.gaccel = E(ga);
        mi.daccel = E(da);
        mi.dtorque = E(dq);
        mi.min_distance = mind;
        mi.max_influence = maxi;
    }
}

}  // namespace

void cpu_deform_by(const CpuSystem &s, int64_t i) {
    eph_mass &mi = s.m[i];
    const DeformSum sum = deform_sum(s, i, tidal_class(mi.k2, mi.tide_delay), V(mi.w));
    m3 C, GI;
    harmonics_of(mi).apply(sum.C, V(mi.w), C, GI);              // UPDATE_HARMONICS
    mi.phi = sum.phi;
    mi.naccel = E(sum.nacc);
    mi.C_potential = E(C);
    mi.GI = E(GI);
}

void cpu_deform(const CpuSystem &s, WorkerPool *pool) {
    // deform_by of one body reads only r, v, GM, sR2 and tide_delay_factor
    // of the others, which deform never writes: the bodies are independent
    run(pool, s.n, [&](int64_t i) { cpu_deform_by(s, i); });
}

void cpu_accel(const CpuSystem &s, int parallel_option, WorkerPool *pool) {
    const int64_t mn = s.n;
    constexpr int64_t n_tasklimit = 64, n_tasksize = 48 * 48;
    int64_t n_task = (int64_t)parallel_option | std::min(mn, (mn * mn + (n_tasksize - 1)) / n_tasksize);
    if (n_task > 1 && pool) {                                   // thread-pool branch, mass.cpp:231-307
        n_task = std::min(n_task, n_tasklimit);
        pool->run(mn, [&](int64_t i) { deform_spin_prepare(s, i); });
        std::vector<JMain> J(mn);
        for (int64_t j = 0; j < mn; ++j) J[j] = jmain_of(s.m[j]);
        pool->run(mn, [&](int64_t i) {                          // accel_mainforce, mass.cpp:136-157
            eph_mass &mi = s.m[i];
            const IMain I = imain_of(mi);
            MainAcc acc;
            acc.mind = mi.min_distance;
            acc.maxi = mi.max_influence;
            for (int64_t j = 0; j < mn; ++j)
                if (j != i) acc.add(pair_terms(I, J[j]));
            mi.gaccel = E(acc.ga);
            mi.daccel = E(acc.da);
            mi.dtorque = E(acc.dq);
            mi.min_distance = acc.mind;
            mi.max_influence = acc.maxi;
        });
        nonpoint_rounds(s, n_task, pool);
    } else {                                                    // single-thread branch, mass.cpp:309-374
        for (int64_t i = 0; i < mn; ++i) deform_spin_prepare(s, i);
        std::vector<JMain> J(mn);
        for (int64_t j = 0; j < mn; ++j) J[j] = jmain_of(s.m[j]);
        main_serial(s, J);
    }
    if (s.n_tidal_children > 0) {                               // tidal-matrix correction, mass.cpp:376-387
        eph_mass &mp = s.m[s.tidal_parent];
        for (int64_t q = 0; q < s.n_tidal_children; ++q) {
            eph_mass &mc = s.m[s.tidal_children[q]];
            const v3 r = diff(DV(mc.r), DV(mp.r));
            const v3 dg = (s.tidal_matrix * r) / (mc.GM + mp.GM);
            mc.daccel = E(V(mc.daccel) + mp.GM * dg);
            mp.daccel = E(V(mp.daccel) - mc.GM * dg);
        }
    }
    for (int64_t i = 0; i < mn; ++i) {                          // mass.cpp:389-396
        eph_mass &mi = s.m[i];
        mi.phi *= kC2;                                          // FINALIZE_RELATIVITY
        const v3 nacc = V(mi.naccel) * kC2;
        mi.naccel = E(nacc);
        if (mi.ring >= 0) {                                     // RING_CORRECTION
            v3 da = V(mi.daccel), dq = V(mi.dtorque);
            ring_correction(s.rings[mi.ring], mi.R2, todbl(DV(mi.GL)), V(mi.gaccel), nacc, da, dq);
            mi.daccel = E(da);
            mi.dtorque = E(dq);
        }
    }
}

void cpu_update(const CpuSystem &s, double t) {
    for (int64_t i = 0; i < s.n; ++i) {
        eph_mass &mi = s.m[i];
        mi.GM = mi.GM0 + mi.dGM * t;
        mi.exJ2 = mi.dJ2 * t;
    }
}

namespace {

struct State0 { ddv3 r, v, GL; ddm3 s; v3 w; };       // mass_state of the step start
struct Stage { ddv3 v; v3 a, w, dtorque; };           // f[k] of RungeKutta.impl:121-127

// RungeKutta12 (RungeKutta.impl:107-194) with the CPU accel after every stage
void rk12(const CpuSystem &s, dd t_eph, double dt, int64_t n_step, WorkerPool *pool) {
    const int64_t mn = s.n;
    std::vector<State0> x0(mn);
    std::vector<Stage> f((size_t)(25 * mn));
    std::vector<ddv3> cv(mn);
    std::vector<v3> ca(mn), cw(mn), ct(mn);
    for (int64_t i_step = 0; i_step < n_step; ++i_step) {
        for (int64_t i = 0; i < mn; ++i) {
            eph_mass &xi = s.m[i];
            x0[i] = State0{DV(xi.r), DV(xi.v), DV(xi.GL), DM(xi.s), V(xi.w)};
            xi.Egrad = xi.dtorque;
        }
        for (int k = 1; k <= 25; ++k) {
            double dt_k = 0;
            for (int64_t i = 0; i < mn; ++i) {
                eph_mass &xi = s.m[i];
                Stage &fi = f[(size_t)(mn * (k - 1) + i)];
                fi.v = DV(xi.v);
                fi.a = (V(xi.gaccel) + V(xi.daccel)) + V(xi.naccel);
                fi.w = V(xi.w);
                fi.dtorque = V(xi.dtorque);
                if (k == 25)
                    xi.Egrad = E(egrad_k25(M(xi.GI), todbl(DV(xi.GL)), V(xi.w), V(xi.dtorque), V(xi.Egrad), dt));
                cv[i] = mkddv(v3{0, 0, 0});
                ca[i] = cw[i] = ct[i] = v3{0, 0, 0};
            }
            for (int j = 0; j < k; ++j) {
                const double *c = kRKTable[k * (k - 1) / 2 + j];
                const dd ckj{c[0], c[1]};
                const double fckj = c[0] + c[1];
                if (ckj.hi == 0) continue;
                dt_k += fckj;
                for (int64_t i = 0; i < mn; ++i) {
                    const Stage &fi = f[(size_t)(mn * j + i)];
                    cv[i] = add(cv[i], scale(ckj, fi.v));
                    ca[i] += fckj * fi.a;
                    cw[i] += fckj * fi.w;
                    ct[i] += fckj * fi.dtorque;
                }
            }
            const double t = todbl(t_eph) + dt * (dt_k + (double)i_step);
            run(pool, mn, [&](int64_t i) {
                const State0 &xi0 = x0[i];
                eph_mass &xi = s.m[i];
                xi.r = E(add(xi0.r, scale(mkdd(dt), cv[i])));
                xi.v = E(add(xi0.v, mkddv(dt * ca[i])));
                const v3 w = cw[i] / dt_k;
                ddm3 sm = xi0.s;
                const v3 dw = (2.0 / 3) * (w - xi0.w);
                sm = add(sm, rotation_matrix(w - dw, dt * dt_k / 2, HostSin{}) * todbl(sm));
                sm = add(sm, rotation_matrix(w + dw, dt * dt_k / 2, HostSin{}) * todbl(sm));
                xi.GL = E(add(xi0.GL, mkddv(dt * ct[i])));
                if (k == 25) {
                    sm = add(sm, rotation_matrix(V(xi.Egrad), 1.0 / 72, HostSin{}) * todbl(sm));
                    sm = orthogonalized(sm);
                }
                xi.s = E(sm);
                xi.w = E(w);
                xi.naccel = E(ca[i]);
                xi.dtorque = E(ct[i]);
                xi.GM = xi.GM0 + xi.dGM * t;                    // msystem::update
                xi.exJ2 = xi.dJ2 * t;
            });
            cpu_accel(s, 0, pool);
        }
    }
}

}  // namespace

void cpu_integrate(const CpuSystem &s, dd &t_eph, double dt, int64_t n_step, WorkerPool *pool) {
    if (n_step == 0) return;
    if (n_step < 0) {
        n_step = -n_step;
        dt = -dt;
    }
    for (int64_t i = 0; i < s.n; ++i) s.m[i].max_influence = s.m[i].min_distance = 0;
    rk12(s, t_eph, dt, n_step, pool);
    for (int64_t i = 0; i < s.n; ++i) s.m[i].min_distance = 1 / s.m[i].min_distance;
    t_eph = add(t_eph, mkdd(dt * (double)n_step));
}

}  // namespace eph
