// Host runtime of the C ABI declared in include/ephemeris_hip.h: batches of
// systems resident in HBM, launch plans for both summation orders, the RK12
// driver loop and HIP-event kernel timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "eph_kernels.hpp"
#include "eph_runtime.hpp"

namespace eph {

static thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}


static const double kRK[325][2] = {
#include "../tables/rk12_feagin.inc"
};

// host double-double add, src/math/dfloat_t.h:83-89
static dd host_add(dd x, dd y) { return add(x, y); }

// ---- eph_mass <-> field-major transposition ----------------------------------
namespace {
#define EM(member) offsetof(eph_mass, member)
std::vector<FieldMap> mass_fields() {
    std::vector<FieldMap> m;
    auto ddv = [&](int f, size_t o) { for (int q = 0; q < 6; ++q) m.push_back({f + q, o + 8 * q}); };
    auto v = [&](int f, size_t o, int k) { for (int q = 0; q < k; ++q) m.push_back({f + q, o + 8 * q}); };
    ddv(F_R, EM(r)); ddv(F_V, EM(v)); ddv(F_GL, EM(GL));
    v(F_S, EM(s), 18); v(F_W, EM(w), 3);
    v(F_GM, EM(GM), 1); v(F_EXJ2, EM(exJ2), 1);
    v(F_GM0, EM(GM0), 1); v(F_DGM, EM(dGM), 1); v(F_DJ2, EM(dJ2), 1); v(F_RAD, EM(R), 1);
    v(F_INERTIA, EM(inertia), 1); v(F_K2, EM(k2), 1); v(F_K2R, EM(k2r), 1); v(F_TD, EM(tide_delay), 1);
    v(F_TDF, EM(tide_delay_factor), 1); v(F_LUM, EM(lum), 1); v(F_RECPT, EM(recpt), 1);
    v(F_A, EM(A), 1); v(F_R2, EM(R2), 1); v(F_SR2, EM(sR2), 1); v(F_RR2G4C, EM(rR2G_4c), 1);
    v(F_CST, EM(C_static), 9);
    v(F_PHI, EM(phi), 1); v(F_BETA2, EM(beta2), 1);
    v(F_NACC, EM(naccel), 3); v(F_DACC, EM(daccel), 3); v(F_DTQ, EM(dtorque), 3);
    v(F_GACC, EM(gaccel), 3); v(F_BETA, EM(beta), 3);
    v(F_GI, EM(GI), 9); v(F_CPOT, EM(C_potential), 9);
    v(F_EGRAD, EM(Egrad), 3); v(F_IDACC, EM(idaccel), 3); v(F_IDTQ, EM(idtorque), 3);
    v(F_JDACC, EM(jdaccel), 3); v(F_JDTQ, EM(jdtorque), 3);
    v(F_MIND, EM(min_distance), 1); v(F_MAXI, EM(max_influence), 1);
    return m;
}
#undef EM
const std::vector<FieldMap> &fields() {
    static const std::vector<FieldMap> f = mass_fields();
    return f;
}
}  // namespace

const std::vector<FieldMap> &mass_field_map() { return fields(); }

// ---- profiling --------------------------------------------------------------
const char *const kKernelNames[kNumKernelClasses] = {
    "stage_update", "deform_spin", "main_force_pool", "main_force_serial", "nonpoint", "small_system_steps",
    "particle_steps"};

bool Profiler::begin(int cls, hipStream_t st) {
    if (!enabled || !((mask >> cls) & 1u)) return false;
    Rec r;
    r.cls = cls;
    if (free_ev.size() < 2) {
        hipEvent_t a, b;
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        free_ev.push_back(a);
        free_ev.push_back(b);
    }
    r.a = free_ev.back(); free_ev.pop_back();
    r.b = free_ev.back(); free_ev.pop_back();
    (void)hipEventRecord(r.a, st);
    open.push_back(r);
    return true;
}
void Profiler::end(hipStream_t st) {
    if (!enabled || open.empty()) return;
    (void)hipEventRecord(open.back().b, st);
    pending.push_back(open.back());
    open.pop_back();
}
void Profiler::collect() {
    for (auto &r : pending) {
        float ms = 0;
        (void)hipEventSynchronize(r.b);
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        total_ms[r.cls] += ms;
        count[r.cls] += 1;
        free_ev.push_back(r.a);
        free_ev.push_back(r.b);
    }
    pending.clear();
}
void Profiler::reset() {
    collect();
    for (int i = 0; i < kNumKernelClasses; ++i) { total_ms[i] = 0; count[i] = 0; }
}
Profiler::~Profiler() {
    for (auto e : free_ev) (void)hipEventDestroy(e);
}

struct Scope {
    Profiler &p;
    hipStream_t st;
    bool on;
    Scope(Profiler &p_, int cls, hipStream_t s) : p(p_), st(s), on(p_.begin(cls, s)) {}
    ~Scope() { if (on) p.end(st); }
};

// ---- system ------------------------------------------------------------------
Batch::~Batch() {
    drop_graph();
    profiler.collect();
    for (void *p : dev_allocs) (void)hipFree(p);
    for (void *p : class_allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
}

template <typename T>
static int dev_upload(Batch &b, T *&dst, const std::vector<T> &src) {
    dst = nullptr;
    if (src.empty()) return EPH_OK;
    HIP_CHECK(hipMalloc(&dst, sizeof(T) * src.size()));
    b.dev_allocs.push_back(dst);
    HIP_CHECK(hipMemcpy(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return EPH_OK;
}

static void append_blocks(std::vector<int2> &blk, int s, int n) {
    for (int i0 = 0; i0 < n; i0 += kBS) blk.push_back(make_int2(s, i0));
}

// Workgroup b of a launch is dispatched to XCD b % 8, and each XCD has its
// own L2.  A work list of many systems is dealt so that every workgroup of
// one system runs on one XCD: that L2 then holds the system's partner data
// and no other XCD fetches it.  Systems go, largest first, to the XCD with
// the least work so far; position 8 r + x of the result holds the r-th item
// of XCD x, and shorter shares are padded with `empty` (which the kernels
// skip).  Lists of few systems, or with one system of more than an eighth of
// the work, are left as they are: spreading each system over every XCD then
// balances better.  Only the placement changes, never any result.
static constexpr int kXcds = 8;

template <typename T, typename SysOf, typename Cost>
static std::vector<T> xcd_deal(const std::vector<T> &items, int64_t n_sys, SysOf sys_of, Cost cost, T empty) {
    std::vector<std::vector<int>> by_sys(n_sys);
    std::vector<double> w(n_sys, 0.0);
    double total = 0, wmax = 0;
    for (int i = 0; i < (int)items.size(); ++i) {
        const int s = sys_of(items[i]);
        by_sys[s].push_back(i);
        w[s] += cost(items[i]);
    }
    std::vector<int> order;
    for (int64_t s = 0; s < n_sys; ++s)
        if (!by_sys[s].empty()) {
            order.push_back((int)s);
            total += w[s];
            wmax = std::max(wmax, w[s]);
        }
    if ((int)order.size() < 2 * kXcds || wmax * kXcds > total) return items;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] > w[b]; });
    std::array<std::vector<T>, kXcds> share;
    std::array<double, kXcds> load{};
    for (int s : order) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += w[s];
        for (int i : by_sys[s]) share[x].push_back(items[i]);
    }
    size_t len = 0;
    for (const auto &sh : share) len = std::max(len, sh.size());
    std::vector<T> out(len * kXcds, empty);
    for (int x = 0; x < kXcds; ++x)
        for (size_t r = 0; r < share[x].size(); ++r) out[r * kXcds + x] = share[x][r];
    return out;
}

int Batch::build_plan(int mode) {
    Plan &pl = plan[mode];
    std::vector<SysMeta> meta(n_sys);
    std::vector<NPItem> items;
    std::vector<int2> blk_pool, blk_serial, blk_recv, work_pool;
    int64_t an_rows = 0;
    std::vector<NPItem> sitems;
    std::vector<int32_t> srow_gp, srow_ring;
    int64_t san_rows = 0;
    pl.maxT = 1;
    pl.max_n = 1;
    pl.np_fusable = true;
    for (int64_t s = 0; s < n_sys; ++s) {
        SysMeta &m = meta[s];
        const int64_t n = sys_n[s];
        m.off = (int32_t)sys_off[s];
        m.n = (int32_t)n;
        m.item0 = (int32_t)items.size();
        m.nitem = 0;
        m.T = 0;
        pl.max_n = std::max<int>(pl.max_n, (int)n);
        // msystem::accel task count, mass.cpp:226-231
        int64_t sq = (n * n + 2303) / 2304;
        int64_t n_task = std::min<int64_t>(n, sq);
        m.serial = (mode == 1 || n_task <= 1) ? 1 : 0;
        if (m.serial) {
            append_blocks(blk_serial, (int)s, (int)n);
            // each model's field at every partner, formed in parallel before
            // the main force (the single-thread kernel reads it per partner)
            for (int64_t i = 0; i < n; ++i) {
                const int64_t g = sys_off[s] + i;
                for (int kind = 0; kind < 2; ++kind) {
                    const int32_t model = kind == 0 ? host_gp[g] : host_ring[g];
                    if (model < 0) continue;
                    if (srow_gp.empty()) { srow_gp.assign((size_t)nb, -1); srow_ring.assign((size_t)nb, -1); }
                    sitems.push_back(NPItem{(int32_t)g, kind, model, 0, (int32_t)s, (int32_t)san_rows});
                    (kind == 0 ? srow_gp : srow_ring)[(size_t)g] = (int32_t)san_rows;
                    san_rows += n;
                }
            }
            continue;
        }
        n_task = std::min<int64_t>(n_task, 64);
        // items in creation order: last body first, geopotential before ring
        std::vector<NPItem> sys_items;
        for (int64_t i = n; i > 0;) {
            --i;
            const int64_t g = sys_off[s] + i;
            if (host_gp[g] >= 0) sys_items.push_back(NPItem{(int32_t)g, 0, host_gp[g], 0, (int32_t)s, 0});
            if (host_ring[g] >= 0) sys_items.push_back(NPItem{(int32_t)g, 1, host_ring[g], 0, (int32_t)s, 0});
        }
        const int64_t nnp = (int64_t)sys_items.size();
        const int64_t T = std::min(nnp, n_task);
        for (int64_t q = 0; q < nnp; ++q) {
            // chain c holds items [c*nnp/T, (c+1)*nnp/T)
            int64_t c = 0;
            while ((c + 1) * nnp / T <= q) ++c;
            sys_items[q].chain = (int32_t)c;
            sys_items[q].an_off = (int32_t)an_rows;
            an_rows += n;
        }
        m.nitem = (int32_t)nnp;
        m.T = (int32_t)T;
        if (n > kNpFusedMaxN_h || nnp > kNpFusedMaxItems_h) pl.np_fusable = false;
        pl.maxT = std::max<int>(pl.maxT, (int)T);
        items.insert(items.end(), sys_items.begin(), sys_items.end());
        append_blocks(blk_pool, (int)s, (int)n);
        if (nnp) append_blocks(blk_recv, (int)s, (int)n);
        for (int64_t i = 0; i < n; ++i) work_pool.push_back(make_int2((int)s, (int)i));
    }
    // 16-body blocks of the four-lanes-per-body kernel, largest systems first
    std::vector<int2> blk16;
    {
        std::vector<int> order;
        for (int64_t s = 0; s < n_sys; ++s)
            if (!meta[s].serial) order.push_back((int)s);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sys_n[a] > sys_n[b]; });
        for (int s : order)
            for (int i0 = 0; i0 < (int)sys_n[s]; i0 += 16) blk16.push_back(make_int2(s, i0));
        blk16 = xcd_deal(blk16, n_sys, [](int2 b) { return b.x; },
                         [&](int2 b) { return (double)sys_n[b.x]; }, make_int2(0, -1));
    }
    int rc;
    if ((rc = dev_upload(*this, pl.d_blk16, blk16))) return rc;
    pl.n_blk16 = (int)blk16.size();
    if ((rc = dev_upload(*this, pl.d_sys, meta))) return rc;
    if ((rc = dev_upload(*this, pl.d_items, items))) return rc;
    if ((rc = dev_upload(*this, pl.d_blk_pool, blk_pool))) return rc;
    if ((rc = dev_upload(*this, pl.d_blk_serial, blk_serial))) return rc;
    if ((rc = dev_upload(*this, pl.d_blk_recv, blk_recv))) return rc;
    if ((rc = dev_upload(*this, pl.d_work_pool, work_pool))) return rc;
    pl.n_work_pool = (int)work_pool.size();
    pl.n_pool = (int)blk_pool.size();
    pl.n_serial = (int)blk_serial.size();
    pl.n_recv = (int)blk_recv.size();
    pl.nrows = (int)items.size();
    pl.an_rows = an_rows;
    pl.n_sitems = (int)sitems.size();
    if (!sitems.empty()) {
        if ((rc = dev_upload(*this, pl.d_sitems, sitems))) return rc;
        if ((rc = dev_upload(*this, pl.d_srow_gp, srow_gp))) return rc;
        if ((rc = dev_upload(*this, pl.d_srow_ring, srow_ring))) return rc;
        HIP_CHECK(hipMalloc(&pl.d_san, sizeof(double) * 3 * san_rows));
        dev_allocs.push_back(pl.d_san);
    }
    if (an_rows) {
        HIP_CHECK(hipMalloc(&pl.d_an, sizeof(double) * 3 * an_rows));
        dev_allocs.push_back(pl.d_an);
        HIP_CHECK(hipMalloc(&pl.d_part, sizeof(double) * 6 * (size_t)pl.nrows * pl.maxT));
        dev_allocs.push_back(pl.d_part);
    }
    return EPH_OK;
}

View Batch::view(int mode) const {
    const Plan &pl = plan[mode];
    View v;
    v.F = F;
    v.nbp = nbp;
    v.nb = (int)nb;
    v.sys = pl.d_sys;
    v.items = pl.d_items;
    v.gps = d_gps;
    v.rings = d_rings;
    v.np_gp = d_np_gp;
    v.np_ring = d_np_ring;
    v.bsys = d_bsys;
    v.perm = d_perm;
    v.srow_gp = pl.d_srow_gp;
    v.srow_ring = pl.d_srow_ring;
    v.san = pl.d_san;
    v.an = pl.d_an;
    v.part = pl.d_part;
    v.t0 = d_t0;
    return v;
}

// The wave-per-body kernels win while the batch is too small to fill the
// GPU with one thread per body (fewer than ~16k bodies in thread-pool order).
static constexpr int kWpbAutoBodies = 16384;

bool Batch::use_wpb(int mode) const {
    const int variant = eff_variant();
    if (variant == 1 || variant == 4) return false;
    if (variant >= 2) return true;
    return plan[mode].n_work_pool > 0 && plan[mode].n_work_pool <= kWpbAutoBodies;
}

int Batch::accel(int mode, bool prepare_only) {
    const int variant = eff_variant();
    const View v = view(mode);
    const Plan &pl = plan[mode];
    const bool wpb = use_wpb(mode);
    // the non-point item fields depend only on the stage's positions and
    // frames: they ride along in extra workgroups of the deform launch
    const int np_rows = prepare_only ? 0 : pl.nrows;
    bool field_done = false;
    {
        Scope sc(profiler, 1, stream);
        if (wpb) {
            launch_deform_wpb(v, d_work_all, n_work_all, max_sys_n, prepare_only ? 0 : 1, np_rows, eff_waves(), stream);
            field_done = true;
        } else if (variant == 1) {
            launch_deform(v, d_blk_light, n_blk_light, d_blk_damped, n_blk_damped, prepare_only ? 0 : 1, stream);
        } else {
            // constant-time-lag bodies re-sum their partners once per spin pass:
            // one wavefront each; the rest four lanes each (that launch
            // carries the item fields)
            launch_deform_mixed(v, d_work_damped, n_work_damped, d_blk_light16, n_blk_light16,
                                prepare_only ? 0 : 1, np_rows, pl.max_n, stream);
            field_done = true;
        }
    }
    if (prepare_only) return EPH_OK;
    const bool fused_np = wpb && pl.np_fusable && variant != 3;
    bool items_done = false;
    if (fused_np && pl.nrows && !field_done) {
        Scope sc(profiler, 4, stream);
        launch_np_field(v, pl.nrows, pl.max_n, stream);
    }
    if (pl.n_pool) {
        Scope sc(profiler, 2, stream);
        if (wpb) launch_main_wpb(v, pl.d_work_pool, pl.n_work_pool, pl.max_n, fused_np, eff_waves(), stream);
        else if (variant == 1) launch_main_pool(v, pl.d_blk_pool, pl.n_pool, stream);
        else if (pl.nrows && !fused_np) {                      // the items' sums ride along
            launch_main_k4_items(v, pl.d_blk16, pl.n_blk16, pl.nrows, pl.maxT, stream);
            items_done = true;
        } else launch_main_k4(v, pl.d_blk16, pl.n_blk16, stream);
    }
    if (pl.n_serial) {
        Scope sc(profiler, 3, stream);
        if (pl.n_sitems) {                                 // the models' fields at every partner
            View vs = v;
            vs.items = pl.d_sitems;
            vs.an = pl.d_san;
            launch_np_field(vs, pl.n_sitems, pl.max_n, stream);
        }
        launch_main_serial(v, pl.d_blk_serial, pl.n_serial, stream);
    }
    if (pl.nrows && !fused_np) {
        Scope sc(profiler, 4, stream);
        launch_np(v, pl.nrows, pl.max_n, pl.maxT, pl.d_blk_recv, pl.n_recv, !field_done, stream, !items_done);
    }
    return EPH_OK;
}

DtK stage_dtk() {                                    // RungeKutta.impl:148-152
    DtK d;
    d.v[0] = 0;
    for (int k = 1; k <= 25; ++k) {
        double dt_k = 0;
        for (int j = 0; j < k; ++j) {
            const double *c = kRK[k * (k - 1) / 2 + j];
            if (c[0] != 0) dt_k += c[0] + c[1];
        }
        d.v[k] = dt_k;
    }
    return d;
}

bool Batch::all_small() const {
    return plan[0].n_pool == 0 && plan[0].n_serial > 0 &&
           *std::max_element(sys_n.begin(), sys_n.end()) <= kSmallMax;
}

// One launch runs many steps for every system (one workgroup per system).
// With `rec`, the field records of each chunk of rec_steps steps are written
// and after_chunk(ctx, first_step, steps) runs before the next chunk.
int Batch::integrate_small(double dt, int64_t n_step, JMain *rec, int64_t rec_steps,
                           void (*after_chunk)(void *, int64_t, int64_t), void *ctx) {
    if (n_step == 0) return EPH_OK;
    if (n_step < 0) { n_step = -n_step; dt = -dt; }
    std::vector<double> t0(n_sys);
    for (int64_t s = 0; s < n_sys; ++s) t0[s] = todbl(t_eph[s]);
    HIP_CHECK(hipMemcpyAsync(d_t0, t0.data(), sizeof(double) * n_sys, hipMemcpyHostToDevice, stream));
    const View v = view(0);
    const DtK dtk = stage_dtk();
    launch_extrema(v, 0, stream);
    const int64_t chunk = rec ? rec_steps : std::min<int64_t>(n_step, 64);
    for (int64_t s0 = 0; s0 < n_step; s0 += chunk) {
        const int64_t ns = std::min(chunk, n_step - s0);
        {
            Scope sc(profiler, 5, stream);
            launch_small(v, d_small_ids, (int)small_ids.size(), max_sys_n, (int)ns, dt, dtk, s0, rec, stream);
        }
        if (after_chunk) after_chunk(ctx, s0, ns);
    }
    launch_extrema(v, 1, stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(stream));
    for (int64_t s = 0; s < n_sys; ++s) t_eph[s] = host_add(t_eph[s], mkdd(dt * (double)n_step));
    return EPH_OK;
}

void Batch::drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
}

// One RK12 step (25 stage updates and force evaluations, then the step
// counter) captured as a graph, rebuilt when dt, the summation order or the
// kernel choice changes.
int Batch::step_graph(double dt, int mode) {
    if (gexec && g_dt == dt && g_mode == mode && g_variant == eff_variant() && g_waves == eff_waves()) return EPH_OK;
    drop_graph();
    if (!d_istep) {
        HIP_CHECK(hipMalloc(&d_istep, sizeof(int64_t)));
        dev_allocs.push_back(d_istep);
    }
    const View v = view(mode);
    const DtK dtk = stage_dtk();
    const bool prof = profiler.enabled;
    profiler.enabled = false;                          // no event records inside the capture
    HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = EPH_OK;
    for (int k = 1; k <= 25 && !rc; ++k) {
        launch_stage(v, k, dt, dtk.v[k], 0.0, d_istep, stream);
        rc = accel(mode, false);
    }
    launch_step_count(d_istep, stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &g);
    profiler.enabled = prof;
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(EPH_EDEVICE, std::string("step graph capture failed: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) { gexec = nullptr; return fail(EPH_EDEVICE, "step graph instantiation failed"); }
    g_dt = dt;
    g_mode = mode;
    g_variant = eff_variant();
    g_waves = eff_waves();
    return EPH_OK;
}

// ---- autotuner ------------------------------------------------------------------
// For a batch in automatic mode with enough thread-pool-order bodies, time
// the block shapes of the force passes on a device copy of the batch --
// one thread per body (64-body blocks, 64-partner j-tiles), one workgroup
// of W = 1, 2 or 4 wavefronts per body (j-tiles of 64 W partners per
// round), four lanes per body (16 bodies per wavefront, 4-partner groups)
// -- and keep the fastest (results are bit-identical, so only speed
// differs).  W = 2 and 4 are tried only while their workgroups can still
// share the GPU in few rounds; choices are cached per shape bucket for the
// life of the process.
namespace {
constexpr int kTuneMinBodies = 2048;
std::mutex g_tune_mu;
std::map<std::array<int, 4>, std::pair<int, int>> g_tune_cache;
int log2_bucket(int64_t x) { int b = 0; while (x > 1) { x >>= 1; ++b; } return b; }
}  // namespace

int Batch::autotune(double dt, int mode) {
    tune_done = true;
    const Plan &pl = plan[mode];
    if (pl.n_work_pool < kTuneMinBodies) return EPH_OK;
    const std::array<int, 4> key{log2_bucket(pl.n_work_pool), (pl.max_n + 63) / 64, log2_bucket(n_sys),
                                 pl.nrows > 0 ? 1 : 0};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            tuned = it->second.first;
            tuned_waves = it->second.second;
            return EPH_OK;
        }
    }
    std::vector<int> all((size_t)n_sys);
    for (int64_t k = 0; k < n_sys; ++k) all[k] = (int)k;
    std::vector<std::pair<int, int>> cands{{1, 0}, {2, 1}};   // (variant, wavefronts per body)
    if (pl.n_work_pool <= 8192) cands.push_back({2, 2});
    if (pl.n_work_pool <= 4096) cands.push_back({2, 4});
    cands.push_back({4, 0});
    std::pair<int, int> best{0, 0};
    double best_s = 1e300;
    for (const auto &cand : cands) {
        Batch *c = nullptr;
        if (regroup_batch(this, all, 0, nullptr, nullptr, nullptr, 0, nullptr, 0, nullptr, &c) || !c) {
            g_last_error.clear();                          // not enough memory for a copy: keep the heuristic
            return EPH_OK;
        }
        std::unique_ptr<Batch> own(c);
        c->variant = cand.first;
        c->waves = cand.second;
        c->use_graph = use_graph;
        c->force_generic = true;
        c->tune_done = true;
        // warm-up of two steps: a graph is captured only for n_step >= 2, so
        // capture and instantiation happen here, outside the timed calls
        int rc = c->integrate(dt, 2, mode);
        double el = 1e300;
        for (int rep = 0; rep < 3 && !rc; ++rep) {         // best of three two-step replays
            const auto t0 = std::chrono::steady_clock::now();
            rc = c->integrate(dt, 2, mode);
            el = std::min(el, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
        if (rc) return rc;
        if (el < best_s) { best_s = el; best = cand; }
    }
    tuned = best.first;
    tuned_waves = best.second;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
    return EPH_OK;
}

int Batch::integrate(double dt, int64_t n_step, int mode) {
    if (n_step == 0) return EPH_OK;
    if ((mode == 1 || plan[0].n_pool == 0) && all_small() && !force_generic)
        return integrate_small(dt, n_step, nullptr, 0, nullptr, nullptr);
    if (variant == 0 && use_autotune && !tune_done) {
        int rc = autotune(n_step < 0 ? -dt : dt, mode);
        if (rc) return rc;
    }
    if (n_step < 0) { n_step = -n_step; dt = -dt; }
    std::vector<double> t0(n_sys);
    for (int64_t s = 0; s < n_sys; ++s) t0[s] = todbl(t_eph[s]);
    HIP_CHECK(hipMemcpyAsync(d_t0, t0.data(), sizeof(double) * n_sys, hipMemcpyHostToDevice, stream));
    const View v = view(mode);
    launch_extrema(v, 0, stream);
    if (use_graph && !profiler.enabled && n_step >= 2) {
        int rc = step_graph(dt, mode);
        if (rc) return rc;
        const int64_t zero = 0;
        HIP_CHECK(hipMemcpyAsync(d_istep, &zero, sizeof zero, hipMemcpyHostToDevice, stream));
        for (int64_t i_step = 0; i_step < n_step; ++i_step) HIP_CHECK(hipGraphLaunch(gexec, stream));
        launch_extrema(v, 1, stream);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(stream));
        for (int64_t s = 0; s < n_sys; ++s) t_eph[s] = host_add(t_eph[s], mkdd(dt * (double)n_step));
        return EPH_OK;
    }
    for (int64_t i_step = 0; i_step < n_step; ++i_step) {
        for (int k = 1; k <= 25; ++k) {
            double dt_k = 0;                                   // RungeKutta.impl:148-152
            for (int j = 0; j < k; ++j) {
                const double *c = kRK[k * (k - 1) / 2 + j];
                if (c[0] != 0) dt_k += c[0] + c[1];
            }
            const double tadd = dt * (dt_k + (double)i_step);
            {
                Scope sc(profiler, 0, stream);
                launch_stage(v, k, dt, dt_k, tadd, nullptr, stream);
            }
            int rc = accel(mode, false);
            if (rc) return rc;
        }
        if (profiler.enabled && (i_step & 15) == 15) profiler.collect();
    }
    launch_extrema(v, 1, stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(stream));
    for (int64_t s = 0; s < n_sys; ++s) t_eph[s] = host_add(t_eph[s], mkdd(dt * (double)n_step));
    return EPH_OK;
}

// Slot permutation per system: bodies with delayed tides (k2 != 0 and
// tide_delay != 0) first, then the rest, each in index order; the deform
// kernels take one class per workgroup.  Partner order is unaffected.
int Batch::build_class_tables(const eph_mass *masses) {
    std::vector<int32_t> perm(nb);
    std::vector<int4> light, damped;
    std::vector<int2> damped_work;       // (system, body) of every constant-time-lag body
    std::vector<int4> light16;           // 16-body blocks of the other bodies
    for (int64_t s = 0; s < n_sys; ++s) {
        int64_t k = sys_off[s];
        for (int pass = 0; pass < 2; ++pass)
            for (int64_t g = sys_off[s]; g < sys_off[s + 1]; ++g) {
                const bool dmp = masses[g].k2 != 0 && masses[g].tide_delay != 0;
                if (dmp == (pass == 0)) perm[k++] = (int32_t)g;
            }
        int64_t nd = 0;
        for (int64_t g = sys_off[s]; g < sys_off[s + 1]; ++g)
            if (masses[g].k2 != 0 && masses[g].tide_delay != 0) {
                ++nd;
                damped_work.push_back(make_int2((int)s, (int)(g - sys_off[s])));
            }
        for (int64_t q = 0; q < nd; q += kBS) damped.push_back(make_int4((int)s, (int)q, (int)std::min<int64_t>(kBS, nd - q), 0));
        for (int64_t q = nd; q < sys_n[s]; q += kBS)
            light.push_back(make_int4((int)s, (int)q, (int)std::min<int64_t>(kBS, sys_n[s] - q), 0));
        for (int64_t q = nd; q < sys_n[s]; q += 16)
            light16.push_back(make_int4((int)s, (int)q, (int)std::min<int64_t>(16, sys_n[s] - q), 0));
    }
    // constant-time-lag bodies re-sum their partners once per spin pass
    damped_work = xcd_deal(damped_work, n_sys, [](int2 b) { return b.x; },
                           [&](int2 b) { return (double)sys_n[b.x]; }, make_int2(0, -1));
    light16 = xcd_deal(light16, n_sys, [](int4 b) { return b.x; },
                       [&](int4 b) { return (double)b.z * sys_n[b.x]; }, make_int4(0, 0, 0, 0));
    // the merged deform launch puts the block list after this one: a length
    // that is a multiple of 8 keeps each block on the XCD it was dealt to
    while (damped_work.size() % kXcds) damped_work.push_back(make_int2(0, -1));
    HIP_CHECK(hipStreamSynchronize(stream));
    for (void *p : class_allocs) (void)hipFree(p);
    class_allocs.clear();
    auto up = [&](auto *&dst, const auto &src) -> int {
        dst = nullptr;
        if (src.empty()) return EPH_OK;
        HIP_CHECK(hipMalloc(&dst, sizeof(src[0]) * src.size()));
        class_allocs.push_back(dst);
        HIP_CHECK(hipMemcpy(dst, src.data(), sizeof(src[0]) * src.size(), hipMemcpyHostToDevice));
        return EPH_OK;
    };
    int rc;
    if ((rc = up(d_perm, perm))) return rc;
    if ((rc = up(d_blk_light, light))) return rc;
    if ((rc = up(d_blk_damped, damped))) return rc;
    if ((rc = up(d_work_damped, damped_work))) return rc;
    if ((rc = up(d_blk_light16, light16))) return rc;
    n_blk_light16 = (int)light16.size();
    n_blk_light = (int)light.size();
    n_blk_damped = (int)damped.size();
    n_work_damped = (int)damped_work.size();
    return EPH_OK;
}

int Batch::upload(const eph_mass *masses) {
    drop_graph();                        // the class tables are rebuilt below
    host_k2.resize(nb);
    host_td.resize(nb);
    for (int64_t g = 0; g < nb; ++g) { host_k2[g] = masses[g].k2; host_td[g] = masses[g].tide_delay; }
    {
        int rc = build_class_tables(masses);
        if (rc) return rc;
    }
    std::vector<double> buf((size_t)kMassFieldEnd * nbp, 0.0);
    for (int64_t g = 0; g < nb; ++g) {
        const char *m = (const char *)&masses[g];
        for (const auto &f : fields()) {
            double x;
            memcpy(&x, m + f.offset, 8);
            buf[(size_t)f.field * nbp + g] = x;
        }
        buf[(size_t)F_PHIC * nbp + g] = masses[g].phi / kC2;
        buf[(size_t)F_NACCC * nbp + g] = masses[g].naccel.x * (1 / kC2);
        buf[(size_t)(F_NACCC + 1) * nbp + g] = masses[g].naccel.y * (1 / kC2);
        buf[(size_t)(F_NACCC + 2) * nbp + g] = masses[g].naccel.z * (1 / kC2);
    }
    HIP_CHECK(hipMemcpy(F, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
    sid.resize(nb);
    for (int64_t g = 0; g < nb; ++g) sid[g] = masses[g].sid;
    return EPH_OK;
}

int Batch::download(eph_mass *masses) {
    std::vector<double> buf((size_t)kMassFieldEnd * nbp);
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipMemcpy(buf.data(), F, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
    for (int64_t g = 0; g < nb; ++g) {
        char *m = (char *)&masses[g];
        memset(m, 0, sizeof(eph_mass));
        for (const auto &f : fields()) {
            double x = buf[(size_t)f.field * nbp + g];
            memcpy(m + f.offset, &x, 8);
        }
        masses[g].sid = sid[g];
        masses[g].gp = host_gp_global[g];
        masses[g].ring = host_ring_global[g];
    }
    return EPH_OK;
}

int create_batch(int64_t n_sys, const int64_t *sys_n, const eph_mass *masses,
                 const eph_geopotential *gps, int64_t n_gp, const eph_ring *rings, int64_t n_ring,
                 const double *t_eph, Batch **out) {
    if (!out || n_sys <= 0 || !sys_n || !masses) return fail(EPH_EINVAL, "eph_system_create: null or empty input");
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    auto b = new Batch();
    b->device = dev;
    b->n_sys = n_sys;
    b->sys_n.assign(sys_n, sys_n + n_sys);
    b->sys_off.resize(n_sys + 1, 0);
    for (int64_t s = 0; s < n_sys; ++s) {
        if (sys_n[s] <= 0 || sys_n[s] > (1 << 24)) { delete b; return fail(EPH_EINVAL, "system size out of range"); }
        b->sys_off[s + 1] = b->sys_off[s] + sys_n[s];
    }
    b->nb = b->sys_off[n_sys];
    b->max_sys_n = *std::max_element(b->sys_n.begin(), b->sys_n.end());
    b->nbp = (b->nb + 63) / 64 * 64;
    b->t_eph.resize(n_sys);
    for (int64_t s = 0; s < n_sys; ++s) b->t_eph[s] = t_eph ? dd{t_eph[2 * s], t_eph[2 * s + 1]} : dd{0, 0};
    b->gps.assign(gps, gps + (gps ? n_gp : 0));
    b->rings.assign(rings, rings + (rings ? n_ring : 0));
    for (const auto &g : b->gps)
        if (g.Nz < 2 || g.Nz > EPH_GP_MAX_DEGREE || g.Nt < 0 || g.Nt > g.Nz) { delete b; return fail(EPH_EINVAL, "bad geopotential degree"); }
    for (const auto &r : b->rings)
        if (r.N < 0 || r.N > EPH_RING_MAX_DISKS) { delete b; return fail(EPH_EINVAL, "bad ring disk count"); }
    b->host_gp.resize(b->nb);
    b->host_ring.resize(b->nb);
    b->host_gp_global.resize(b->nb);
    b->host_ring_global.resize(b->nb);
    std::vector<int32_t> bsys(b->nb);
    for (int64_t s = 0; s < n_sys; ++s)
        for (int64_t g = b->sys_off[s]; g < b->sys_off[s + 1]; ++g) bsys[g] = (int32_t)s;
    for (int64_t g = 0; g < b->nb; ++g) {
        const int64_t gp = masses[g].gp, rg = masses[g].ring;
        if (gp >= n_gp || rg >= n_ring) { delete b; return fail(EPH_EINVAL, "model index out of range"); }
        b->host_gp[g] = gp >= 0 ? (int32_t)gp : -1;
        b->host_ring[g] = rg >= 0 ? (int32_t)rg : -1;
        b->host_gp_global[g] = gp >= 0 ? gp : -1;
        b->host_ring_global[g] = rg >= 0 ? rg : -1;
    }
    int rc;
    auto bail = [&](int code) { delete b; return code; };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(EPH_EDEVICE, "hipStreamCreate failed"));
    size_t bytes = sizeof(double) * (size_t)F_COUNT * b->nbp;
    if (hipMalloc(&b->F, bytes) != hipSuccess) return bail(fail(EPH_ENOMEM, "hipMalloc of the state failed"));
    b->dev_allocs.push_back(b->F);
    if (hipMemset(b->F, 0, bytes) != hipSuccess) return bail(fail(EPH_EDEVICE, "hipMemset failed"));
    if ((rc = dev_upload(*b, b->d_gps, b->gps))) return bail(rc);
    if ((rc = dev_upload(*b, b->d_rings, b->rings))) return bail(rc);
    if ((rc = dev_upload(*b, b->d_np_gp, b->host_gp))) return bail(rc);
    if ((rc = dev_upload(*b, b->d_np_ring, b->host_ring))) return bail(rc);
    if ((rc = dev_upload(*b, b->d_bsys, bsys))) return bail(rc);
    std::vector<double> t0(n_sys, 0.0);
    if ((rc = dev_upload(*b, b->d_t0, t0))) return bail(rc);
    std::vector<int2> blk_all;
    for (int64_t s = 0; s < n_sys; ++s) append_blocks(blk_all, (int)s, (int)sys_n[s]);
    if ((rc = dev_upload(*b, b->d_blk_all, blk_all))) return bail(rc);
    b->n_blk_all = (int)blk_all.size();
    if ((rc = b->build_plan(0))) return bail(rc);
    if ((rc = b->build_plan(1))) return bail(rc);
    for (int64_t s = 0; s < n_sys; ++s) b->small_ids.push_back((int)s);
    {
        std::vector<int2> work_all;
        for (int64_t s = 0; s < n_sys; ++s)
            for (int64_t i = 0; i < sys_n[s]; ++i) work_all.push_back(make_int2((int)s, (int)i));
        if ((rc = dev_upload(*b, b->d_work_all, work_all))) return bail(rc);
        b->n_work_all = (int)work_all.size();
    }
    if ((rc = dev_upload(*b, b->d_small_ids, b->small_ids))) return bail(rc);
    set_rk_table(&kRK[0][0], b->stream);
    if ((rc = b->upload(masses))) return bail(rc);
    if (hipStreamSynchronize(b->stream) != hipSuccess) return bail(fail(EPH_EDEVICE, "stream sync failed"));
    *out = b;
    return EPH_OK;
}


// A new batch holding systems `keep` of `old` (state copied device to device,
// never through the host) followed by n_new systems given on the host.
int regroup_batch(Batch *old, const std::vector<int> &keep, int64_t n_new, const int64_t *new_sys_n,
                  const eph_mass *new_masses, const eph_geopotential *new_gps, int64_t n_new_gp,
                  const eph_ring *new_rings, int64_t n_new_ring, const double *new_t_eph, Batch **out) {
    std::vector<int64_t> sys_n;
    std::vector<double> teph;
    std::vector<eph_geopotential> gps;
    std::vector<eph_ring> rings;
    std::vector<eph_mass> stubs;
    for (int k : keep) {
        const int64_t o = old->sys_off[k], n = old->sys_n[k];
        sys_n.push_back(n);
        teph.push_back(old->t_eph[k].hi);
        teph.push_back(old->t_eph[k].lo);
        for (int64_t g = o; g < o + n; ++g) {
            eph_mass m;
            memset(&m, 0, sizeof m);
            m.k2 = old->host_k2[g];
            m.tide_delay = old->host_td[g];
            m.sid = old->sid[g];
            m.gp = -1;
            m.ring = -1;
            if (old->host_gp_global[g] >= 0) { m.gp = (int64_t)gps.size(); gps.push_back(old->gps[old->host_gp_global[g]]); }
            if (old->host_ring_global[g] >= 0) { m.ring = (int64_t)rings.size(); rings.push_back(old->rings[old->host_ring_global[g]]); }
            stubs.push_back(m);
        }
    }
    const int64_t gp_base = (int64_t)gps.size(), ring_base = (int64_t)rings.size();
    int64_t nn = 0;
    for (int64_t s = 0; s < n_new; ++s) {
        sys_n.push_back(new_sys_n[s]);
        teph.push_back(new_t_eph ? new_t_eph[2 * s] : 0.0);
        teph.push_back(new_t_eph ? new_t_eph[2 * s + 1] : 0.0);
        nn += new_sys_n[s];
    }
    for (int64_t i = 0; i < n_new_gp; ++i) gps.push_back(new_gps[i]);
    for (int64_t i = 0; i < n_new_ring; ++i) rings.push_back(new_rings[i]);
    for (int64_t g = 0; g < nn; ++g) {
        eph_mass m = new_masses[g];
        if (m.gp >= 0) m.gp += gp_base;
        if (m.ring >= 0) m.ring += ring_base;
        stubs.push_back(m);
    }
    if (sys_n.empty()) { *out = nullptr; return EPH_OK; }
    Batch *b = nullptr;
    int rc = create_batch((int64_t)sys_n.size(), sys_n.data(), stubs.data(), gps.data(), (int64_t)gps.size(),
                          rings.data(), (int64_t)rings.size(), teph.data(), &b);
    if (rc) return rc;
    int64_t dst = 0;
    for (int k : keep) {
        const int64_t n = old->sys_n[k];
        if (hipMemcpy2DAsync(b->F + dst, sizeof(double) * b->nbp, old->F + old->sys_off[k], sizeof(double) * old->nbp,
                             sizeof(double) * n, kMassFieldEnd, hipMemcpyDeviceToDevice, b->stream) != hipSuccess) {
            delete b;
            return fail(EPH_EDEVICE, "device-to-device regroup copy failed");
        }
        dst += n;
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) { delete b; return fail(EPH_EDEVICE, "regroup sync failed"); }
    *out = b;
    return EPH_OK;
}


// ---- test particles -------------------------------------------------------------
Particles::~Particles() {
    if (pstream) (void)hipStreamSynchronize(pstream);
    for (int b = 0; b < 2; ++b) {
        if (ev_rec[b]) (void)hipEventDestroy(ev_rec[b]);
        if (ev_free[b]) (void)hipEventDestroy(ev_free[b]);
    }
    if (pstream) (void)hipStreamDestroy(pstream);
    if (P) (void)hipFree(P);
    if (rec) (void)hipFree(rec);
}

// Steps per chunk of field records.  The field computes chunk c + 1 while
// the particles advance through chunk c, so a chunk of one step hides all
// but the first step of the field's integration.
static constexpr int64_t kRecChunk = 1;

// The partner data (JMain) each body of the one-system batch `field`
// presents at its current state -- what the last force evaluation left --
// written to dst[0..kSmallMax) in device memory.
int field_epoch_record(Batch *field, JMain *dst) {
    std::vector<JMain> t0(kSmallMax);
    std::vector<eph_mass> fm(field->nb);
    int rc = field->download(fm.data());
    if (rc) return rc;
    for (int64_t j = 0; j < field->nb; ++j) {
        const eph_mass &m = fm[j];
        JMain &T = t0[j];
        memset(&T, 0, sizeof T);
        T.rxh = m.r.x.hi; T.rxl = m.r.x.lo; T.ryh = m.r.y.hi; T.ryl = m.r.y.lo; T.rzh = m.r.z.hi; T.rzl = m.r.z.lo;
        T.GM = m.GM; T.sR2 = m.sR2;
        T.beta2 = m.beta2;
        T.bx = m.beta.x; T.by = m.beta.y; T.bz = m.beta.z;
        T.R2 = m.R2;
        T.GLx = m.GL.x.hi + m.GL.x.lo; T.GLy = m.GL.y.hi + m.GL.y.lo; T.GLz = m.GL.z.hi + m.GL.z.lo;
        T.lum = m.lum;
        const double *C = &m.C_potential.x.x;
        for (int q = 0; q < 9; ++q) T.C[q] = C[q];
    }
    // phi/c^2 and naccel/c^2 exactly as the device holds them
    const size_t nbp = (size_t)field->nbp;
    std::vector<double> pc(field->nb), nc(3 * field->nb);
    HIP_CHECK(hipMemcpy(pc.data(), field->F + F_PHIC * nbp, sizeof(double) * field->nb, hipMemcpyDeviceToHost));
    for (int c = 0; c < 3; ++c)
        HIP_CHECK(hipMemcpy(nc.data() + c * field->nb, field->F + (F_NACCC + c) * nbp, sizeof(double) * field->nb,
                            hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < field->nb; ++j) {
        t0[j].phi = pc[j];
        t0[j].nx = nc[j]; t0[j].ny = nc[field->nb + j]; t0[j].nz = nc[2 * field->nb + j];
    }
    HIP_CHECK(hipMemcpy(dst, t0.data(), sizeof(JMain) * kSmallMax, hipMemcpyHostToDevice));
    return EPH_OK;
}

int create_particles(Batch *field, int64_t n, const double *r, const double *v, double GM, double R, Particles **out) {
    if (!field || !out || n < 0 || (n && (!r || !v))) return fail(EPH_EINVAL, "eph_particles_create: bad argument");
    if (field->n_sys != 1 || !field->all_small())
        return fail(EPH_EINVAL, "the particle field must be one system of at most 48 bodies (single-thread order)");
    auto *ps = new Particles();
    ps->field = field;
    ps->np = n;
    ps->npp = (n + 255) / 256 * 256;
    if (ps->npp == 0) ps->npp = 256;
    ps->GM = GM;
    ps->R = R;
    std::vector<double> host((size_t)P_COUNT_HOST * ps->npp, 0.0);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            host[(size_t)(2 * c) * ps->npp + i] = r[3 * i + c];
            host[(size_t)(6 + 2 * c) * ps->npp + i] = v[3 * i + c];
        }
    auto bail = [&](int code) { delete ps; return code; };
    if (hipMalloc(&ps->P, sizeof(double) * P_ALLOC_ROWS * ps->npp) != hipSuccess)
        return bail(fail(EPH_ENOMEM, "particle state allocation failed"));
    if (hipMemcpy(ps->P, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(EPH_EDEVICE, "particle upload failed"));
    ps->rec_steps = kRecChunk;
    const size_t rec_bytes = 2 * sizeof(JMain) * kSmallMax * (size_t)(25 * kRecChunk);
    if (hipMalloc(&ps->rec, rec_bytes) != hipSuccess) return bail(fail(EPH_ENOMEM, "field record allocation failed"));
    if (hipStreamCreateWithFlags(&ps->pstream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(EPH_EDEVICE, "hipStreamCreate failed"));
    for (int b = 0; b < 2; ++b)
        if (hipEventCreateWithFlags(&ps->ev_rec[b], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ps->ev_free[b], hipEventDisableTiming) != hipSuccess)
            return bail(fail(EPH_EDEVICE, "hipEventCreate failed"));
    {
        int rc = field_epoch_record(field, ps->rec);
        if (rc) return bail(rc);
    }
    PView pv{ps->P, ps->npp, (int)ps->np, ps->rec, 1, (int)field->nb, R < 0 ? R * R : 0.0};
    launch_particles(pv, -1, 0.0, DtK{}, field->stream);
    if (hipStreamSynchronize(field->stream) != hipSuccess) return bail(fail(EPH_EDEVICE, "particle init failed"));
    *out = ps;
    return EPH_OK;
}

// The field (k_small, on the field's stream) writes the records of chunk c
// into buffer c % 2; the particles (on their own stream) replay them once
// that launch is done; the field reuses a buffer once the particle launch
// that read it is done.  Chunking does not change any result: the particle
// state is stored between launches in full.
int integrate_particles(Particles *ps, double dt, int64_t n_step) {
    if (n_step < 0) return fail(EPH_EINVAL, "particles integrate forward only");
    if (n_step == 0) return EPH_OK;
    Batch *f = ps->field;
    std::vector<double> t0(f->n_sys);
    for (int64_t s = 0; s < f->n_sys; ++s) t0[s] = todbl(f->t_eph[s]);
    HIP_CHECK(hipMemcpyAsync(f->d_t0, t0.data(), sizeof(double) * f->n_sys, hipMemcpyHostToDevice, f->stream));
    const View v = f->view(0);
    const DtK dtk = stage_dtk();
    launch_extrema(v, 0, f->stream);
    const size_t buf = (size_t)25 * ps->rec_steps * kSmallMax;   // records per buffer
    const double sR2 = ps->R < 0 ? ps->R * ps->R : 0.0;
    int64_t c = 0;
    for (int64_t s0 = 0; s0 < n_step; s0 += ps->rec_steps, ++c) {
        const int64_t ns = std::min(ps->rec_steps, n_step - s0);
        const int b = (int)(c & 1);
        JMain *rec = ps->rec + b * buf;
        if (c >= 2) HIP_CHECK(hipStreamWaitEvent(f->stream, ps->ev_free[b], 0));
        {
            Scope sc(f->profiler, 5, f->stream);
            launch_small(v, f->d_small_ids, (int)f->small_ids.size(), f->max_sys_n, (int)ns, dt, dtk, s0, rec, f->stream);
        }
        HIP_CHECK(hipEventRecord(ps->ev_rec[b], f->stream));
        HIP_CHECK(hipStreamWaitEvent(ps->pstream, ps->ev_rec[b], 0));
        {
            PView pv{ps->P, ps->npp, (int)ps->np, rec, 1, (int)f->nb, sR2};
            Scope sc(f->profiler, 6, ps->pstream);
            launch_particles(pv, (int)ns, dt, dtk, ps->pstream);
        }
        HIP_CHECK(hipEventRecord(ps->ev_free[b], ps->pstream));
    }
    HIP_CHECK(hipStreamWaitEvent(f->stream, ps->ev_free[(c - 1) & 1], 0));
    launch_extrema(v, 1, f->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(f->stream));
    for (int64_t s = 0; s < f->n_sys; ++s) f->t_eph[s] = host_add(f->t_eph[s], mkdd(dt * (double)n_step));
    return EPH_OK;
}

int download_particles(Particles *ps, double *r, double *v, double *extrema) {
    std::vector<double> host((size_t)P_COUNT_HOST * ps->npp);
    HIP_CHECK(hipStreamSynchronize(ps->field->stream));
    HIP_CHECK(hipMemcpy(host.data(), ps->P, sizeof(double) * host.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < ps->np; ++i) {
        for (int c = 0; c < 6; ++c) {
            if (r) r[6 * i + c] = host[(size_t)c * ps->npp + i];
            if (v) v[6 * i + c] = host[(size_t)(6 + c) * ps->npp + i];
        }
        if (extrema) {
            extrema[2 * i] = host[(size_t)15 * ps->npp + i];
            extrema[2 * i + 1] = host[(size_t)16 * ps->npp + i];
        }
    }
    return EPH_OK;
}

}  // namespace eph

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using eph::Batch;
using eph::fail;

// eph_system is the opaque C name of a Batch
static Batch *B(eph_system *s) { return reinterpret_cast<Batch *>(s); }
static const Batch *B(const eph_system *s) { return reinterpret_cast<const Batch *>(s); }

extern "C" {

const char *eph_last_error(void) { return eph::g_last_error.c_str(); }
int eph_abi_version(void) { return EPH_ABI_VERSION; }

int eph_init(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(EPH_ENODEV, "no HIP device visible");
    if (device < 0 || device >= n) return fail(EPH_ENODEV, "device index out of range");
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return fail(EPH_ENODEV, "hipGetDeviceProperties failed");
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return fail(EPH_ENODEV, std::string("library built for gfx950, device is ") + p.gcnArchName);
    if (hipSetDevice(device) != hipSuccess) return fail(EPH_ENODEV, "hipSetDevice failed");
    return eph::noalloc_init_device(device);
}

int eph_system_create(int64_t n_sys, const int64_t *sys_n, const eph_mass *masses,
                      const eph_geopotential *gps, int64_t n_gp, const eph_ring *rings, int64_t n_ring,
                      const double *t_eph, eph_system **out) {
    Batch *b = nullptr;
    int rc = eph::create_batch(n_sys, sys_n, masses, gps, n_gp, rings, n_ring, t_eph, &b);
    if (rc) return rc;
    *out = reinterpret_cast<eph_system *>(b);
    return EPH_OK;
}

int eph_system_destroy(eph_system *sys) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    delete B(sys);
    return EPH_OK;
}

int eph_system_upload(eph_system *sys, const eph_mass *masses, const double *t_eph) {
    if (!sys || !masses) return fail(EPH_EINVAL, "null argument");
    int rc = B(sys)->upload(masses);
    if (rc) return rc;
    if (t_eph)
        for (int64_t s = 0; s < B(sys)->n_sys; ++s) B(sys)->t_eph[s] = eph::dd{t_eph[2 * s], t_eph[2 * s + 1]};
    return EPH_OK;
}

int eph_system_download(eph_system *sys, eph_mass *masses, double *t_eph) {
    if (!sys || !masses) return fail(EPH_EINVAL, "null argument");
    int rc = B(sys)->download(masses);
    if (rc) return rc;
    if (t_eph)
        for (int64_t s = 0; s < B(sys)->n_sys; ++s) { t_eph[2 * s] = B(sys)->t_eph[s].hi; t_eph[2 * s + 1] = B(sys)->t_eph[s].lo; }
    return EPH_OK;
}

int64_t eph_system_bodies(const eph_system *sys) { return sys ? B(sys)->nb : -1; }
int64_t eph_system_count(const eph_system *sys) { return sys ? B(sys)->n_sys : -1; }

int eph_deform(eph_system *sys) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    int rc = B(sys)->accel(0, true);
    if (rc) return rc;
    if (hipStreamSynchronize(B(sys)->stream) != hipSuccess) return fail(EPH_EDEVICE, "deform failed");
    return EPH_OK;
}

int eph_accel(eph_system *sys, int summation) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    if (summation != EPH_SUMMATION_POOL && summation != EPH_SUMMATION_SERIAL)
        return fail(EPH_EINVAL, "summation must be EPH_SUMMATION_POOL or EPH_SUMMATION_SERIAL");
    int rc = B(sys)->accel(summation == EPH_SUMMATION_SERIAL ? 1 : 0, false);
    if (rc) return rc;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(EPH_EDEVICE, std::string("accel launch: ") + hipGetErrorString(e));
    if (hipStreamSynchronize(B(sys)->stream) != hipSuccess) return fail(EPH_EDEVICE, "accel failed");
    return EPH_OK;
}

int eph_integrate_ex(eph_system *sys, double dt, int64_t n_step, int summation) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    if (!(dt == dt) || dt == 0) return fail(EPH_EINVAL, "dt must be finite and non-zero");
    if (summation != EPH_SUMMATION_POOL && summation != EPH_SUMMATION_SERIAL)
        return fail(EPH_EINVAL, "summation must be EPH_SUMMATION_POOL or EPH_SUMMATION_SERIAL");
    return B(sys)->integrate(dt, n_step, summation == EPH_SUMMATION_SERIAL ? 1 : 0);
}

int eph_integrate(eph_system *sys, double dt, int64_t n_step) {
    return eph_integrate_ex(sys, dt, n_step, EPH_SUMMATION_POOL);
}

int eph_particles_create(eph_system *field, int64_t n, const double *r, const double *v, double GM, double R,
                         eph_particles **out) {
    eph::Particles *ps = nullptr;
    int rc = eph::create_particles(reinterpret_cast<Batch *>(field), n, r, v, GM, R, &ps);
    if (rc) return rc;
    *out = reinterpret_cast<eph_particles *>(ps);
    return EPH_OK;
}

int eph_particles_integrate(eph_particles *ps, double dt, int64_t n_step) {
    if (!ps) return fail(EPH_EINVAL, "null particles");
    if (!(dt == dt) || dt == 0) return fail(EPH_EINVAL, "dt must be finite and non-zero");
    return eph::integrate_particles(reinterpret_cast<eph::Particles *>(ps), dt, n_step);
}

int eph_particles_download(eph_particles *ps, double *r, double *v, double *extrema) {
    if (!ps) return fail(EPH_EINVAL, "null particles");
    return eph::download_particles(reinterpret_cast<eph::Particles *>(ps), r, v, extrema);
}

int64_t eph_particles_count(const eph_particles *ps) {
    return ps ? reinterpret_cast<const eph::Particles *>(ps)->np : -1;
}

int eph_particles_destroy(eph_particles *ps) {
    if (!ps) return fail(EPH_EINVAL, "null particles");
    delete reinterpret_cast<eph::Particles *>(ps);
    return EPH_OK;
}

int eph_system_set_variant(eph_system *sys, int variant, int allow_persistent) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    if (variant < 0 || variant > 4) return fail(EPH_EINVAL, "variant must be 0 (automatic) or 1-4");
    B(sys)->variant = variant;
    B(sys)->force_generic = !allow_persistent;
    B(sys)->drop_graph();
    return EPH_OK;
}

int eph_system_set_autotune(eph_system *sys, int enable) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    B(sys)->use_autotune = enable != 0;
    return EPH_OK;
}

int eph_system_kernel_variant(const eph_system *sys) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    return B(sys)->eff_variant();
}

int eph_system_set_waves(eph_system *sys, int waves_per_body) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    if (waves_per_body != 0 && waves_per_body != 1 && waves_per_body != 2 && waves_per_body != 4)
        return fail(EPH_EINVAL, "wavefronts per body must be 0 (automatic), 1, 2 or 4");
    B(sys)->waves = waves_per_body;
    B(sys)->drop_graph();
    return EPH_OK;
}

int eph_system_kernel_waves(const eph_system *sys) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    return B(sys)->eff_waves();
}

int eph_system_set_graph(eph_system *sys, int enable) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    B(sys)->use_graph = enable != 0;
    return EPH_OK;
}

int eph_profile_enable(eph_system *sys, int enable) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    B(sys)->profiler.collect();
    B(sys)->profiler.enabled = enable != 0;
    return EPH_OK;
}

int eph_profile_classes(eph_system *sys, unsigned mask) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    B(sys)->profiler.collect();
    B(sys)->profiler.mask = mask;
    return EPH_OK;
}

int eph_profile_reset(eph_system *sys) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    (void)hipStreamSynchronize(B(sys)->stream);
    B(sys)->profiler.reset();
    return EPH_OK;
}

int eph_profile_read(eph_system *sys, const char **names, double *avg_ms, int64_t *launches, int cap) {
    if (!sys) return fail(EPH_EINVAL, "null system");
    (void)hipStreamSynchronize(B(sys)->stream);
    auto &p = B(sys)->profiler;
    p.collect();
    int n = std::min(cap, eph::kNumKernelClasses);
    for (int i = 0; i < n; ++i) {
        if (names) names[i] = eph::kKernelNames[i];
        if (avg_ms) avg_ms[i] = p.count[i] ? p.total_ms[i] / (double)p.count[i] : 0.0;
        if (launches) launches[i] = p.count[i];
    }
    return n;
}

}  // extern "C"
