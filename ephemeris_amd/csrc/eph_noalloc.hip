// Allocation-free force evaluation: eph_accel_noalloc, the variant of
// eph_system_create + eph_accel + eph_system_download for callers that cannot
// allocate device memory.  The caller's bodies and models are in device
// memory; the field-major state, the non-point partial sums and the launch
// plan of the evaluation live in `scratch`, a fixed array of this library on
// each device.  Every call on a device shares that device's array (calls
// that overlap on one device race), and the last call's contents stay.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <mutex>

#include "eph_kernels.hpp"
#include "eph_runtime.hpp"

namespace eph {
namespace {

constexpr int kNaMaxBodies = 1024;     // bodies per call
constexpr int kNaMaxItems = 64;        // geopotential and ring models per call
constexpr int kNaMaxBlocks = kNaMaxBodies / kBS;
constexpr int kNaMaxDevices = 64;

constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }
// scratch layout (bytes)
struct NaPlan {
    SysMeta sys;
    NPItem items[kNaMaxItems];
    NPItem sitems[kNaMaxItems];                                   // single-thread order: each model's field rows
    int32_t srow_gp[kNaMaxBodies], srow_ring[kNaMaxBodies];
    int2 blk_main[kNaMaxBlocks], blk_recv[kNaMaxBlocks];
    int2 work[kNaMaxBodies];                                      // (system 0, body) of every body
    int32_t np_gp[kNaMaxBodies], np_ring[kNaMaxBodies], bsys[kNaMaxBodies];
    double t0;
};
constexpr size_t kOffF = 0;                                                            // fields [0, 121) x 1024
constexpr size_t kOffAn = align256(kOffF + sizeof(double) * kMassFieldEnd * kNaMaxBodies);
constexpr size_t kOffPart = align256(kOffAn + sizeof(double) * 3 * kNaMaxItems * kNaMaxBodies);
constexpr size_t kOffModels = align256(kOffPart + sizeof(double) * 6 * kNaMaxItems * kNaMaxItems);
constexpr size_t kOffPlan = align256(kOffModels + sizeof(int64_t) * 2 * kNaMaxBodies);
constexpr size_t kScratchBytes = align256(kOffPlan + sizeof(NaPlan));     // about 2.8 MB

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// eph_mass member of each field (set at initialisation)
__constant__ int c_na_field[kMassFieldEnd], c_na_off[kMassFieldEnd];
__constant__ int c_na_nmap;

namespace {

std::mutex g_na_mu;
void *g_na_ptr[kNaMaxDevices];        // device address of `scratch`, per device

// bodies (AoS, device) into the field-major state; slots [n, nbp) are zeroed
__global__ void k_na_gather(const eph_mass *__restrict__ in, int n, int64_t nbp, double *__restrict__ F) {
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= nbp) return;
    const bool live = g < n;
    const char *rec = live ? (const char *)(in + g) : nullptr;
    for (int e = 0; e < c_na_nmap; ++e)
        F[(int64_t)c_na_field[e] * nbp + g] = live ? *(const double *)(rec + c_na_off[e]) : 0.0;
    // phi / c^2 and naccel / c^2 as eph_system_upload forms them
    F[(int64_t)F_PHIC * nbp + g] = live ? in[g].phi / kC2 : 0.0;
    F[(int64_t)F_NACCC * nbp + g] = live ? in[g].naccel.x * (1 / kC2) : 0.0;
    F[(int64_t)(F_NACCC + 1) * nbp + g] = live ? in[g].naccel.y * (1 / kC2) : 0.0;
    F[(int64_t)(F_NACCC + 2) * nbp + g] = live ? in[g].naccel.z * (1 / kC2) : 0.0;
}

// the field-major state back into bodies (AoS, device); out may be in
__global__ void k_na_scatter(const double *__restrict__ F, int64_t nbp, int n, const eph_mass *in, eph_mass *out) {
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= n) return;
    const uint64_t sid = in[g].sid;
    const int64_t gp = in[g].gp, ring = in[g].ring;
    char *rec = (char *)(out + g);
    for (int e = 0; e < c_na_nmap; ++e) *(double *)(rec + c_na_off[e]) = F[(int64_t)c_na_field[e] * nbp + g];
    out[g].sid = sid;
    out[g].gp = gp;
    out[g].ring = ring;
}

// the model indices (gp, ring) of every body, contiguous, for one host read
__global__ void k_na_models(const eph_mass *__restrict__ in, int n, int64_t *__restrict__ out) {
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= n) return;
    out[2 * g] = in[g].gp;
    out[2 * g + 1] = in[g].ring;
}

// true when p is device memory of the current device (the kernels read it)
bool on_device(const void *p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();                          // clear the error of a host pointer
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace

int noalloc_init_device(int device) {
    if (device < 0 || device >= kNaMaxDevices) return fail(EPH_ENODEV, "device index beyond the supported count");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (g_na_ptr[device]) return EPH_OK;
    void *ptr = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    const auto &map = mass_field_map();
    int field[kMassFieldEnd], off[kMassFieldEnd];
    const int nmap = (int)map.size();
    for (int e = 0; e < nmap; ++e) {
        field[e] = map[(size_t)e].field;
        off[e] = (int)map[(size_t)e].offset;
    }
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_na_field), field, sizeof(int) * nmap));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_na_off), off, sizeof(int) * nmap));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_na_nmap), &nmap, sizeof nmap));
    g_na_ptr[device] = ptr;
    return EPH_OK;
}

}  // namespace eph

extern "C" int eph_accel_noalloc(const eph_mass *masses, int64_t n, const eph_geopotential *gps, int64_t n_gp,
                                 const eph_ring *rings, int64_t n_ring, int summation, eph_mass *out) {
    using namespace eph;
    if (!masses || !out || n <= 0) return fail(EPH_EINVAL, "eph_accel_noalloc: null or empty argument");
    if (n > kNaMaxBodies) return fail(EPH_EINVAL, "eph_accel_noalloc: more than 1024 bodies");
    if (summation != EPH_SUMMATION_POOL && summation != EPH_SUMMATION_SERIAL)
        return fail(EPH_EINVAL, "summation must be EPH_SUMMATION_POOL or EPH_SUMMATION_SERIAL");
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) return fail(EPH_ENODEV, "device index beyond the supported count");
    if (!g_na_ptr[dev]) {
        int rc = noalloc_init_device(dev);
        if (rc) return rc;
    }
    if (!on_device(masses) || !on_device(out) || (n_gp > 0 && (!gps || !on_device(gps))) ||
        (n_ring > 0 && (!rings || !on_device(rings))))
        return fail(EPH_EINVAL, "eph_accel_noalloc: bodies, models and output must be in device memory");
    char *base = (char *)g_na_ptr[dev];

    // what the plan needs of each body, read from device memory
    int64_t models[2 * kNaMaxBodies];
    hipLaunchKernelGGL(k_na_models, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, masses, (int)n,
                       (int64_t *)(base + kOffModels));
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(models, base + kOffModels, sizeof(int64_t) * 2 * n, hipMemcpyDeviceToHost));
    const int64_t *gp = models;                           // gp[2 i], ring[2 i + 1]

    static thread_local NaPlan plan;                      // host image of the plan (no allocation)
    memset(&plan, 0, sizeof plan);
    int64_t nnp = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t g = gp[2 * i], r = gp[2 * i + 1];
        if (g >= n_gp || r >= n_ring || (g >= 0 && !gps) || (r >= 0 && !rings))
            return fail(EPH_EINVAL, "eph_accel_noalloc: model index out of range");
        plan.np_gp[i] = g >= 0 ? (int32_t)g : -1;
        plan.np_ring[i] = r >= 0 ? (int32_t)r : -1;
        nnp += (g >= 0) + (r >= 0);
    }
    if (nnp > kNaMaxItems) return fail(EPH_EINVAL, "eph_accel_noalloc: more than 64 geopotential and ring models");
    int n_main = 0, n_recv = 0;
    for (int q = 0; q < (int)n; q += kBS) plan.blk_main[n_main++] = make_int2(0, q);
    for (int i = 0; i < (int)n; ++i) plan.work[i] = make_int2(0, i);
    // summation order and the non-point rounds (msystem::accel, mass.cpp:226-307)
    const int64_t n_task = std::min<int64_t>(n, (n * n + 2303) / 2304);
    const bool serial = summation == EPH_SUMMATION_SERIAL || n_task <= 1;
    int n_sitems = 0;
    if (serial) {                                         // each model's field at every partner, formed first
        int32_t san_rows = 0;
        for (int64_t i = 0; i < n; ++i) {
            plan.srow_gp[i] = plan.srow_ring[i] = -1;
            for (int kind = 0; kind < 2; ++kind) {
                const int32_t model = kind == 0 ? plan.np_gp[i] : plan.np_ring[i];
                if (model < 0) continue;
                plan.sitems[n_sitems++] = NPItem{(int32_t)i, kind, model, 0, 0, san_rows};
                (kind == 0 ? plan.srow_gp : plan.srow_ring)[i] = san_rows;
                san_rows += (int32_t)n;
            }
        }
    }
    SysMeta &sm = plan.sys;
    sm.off = 0;
    sm.n = (int32_t)n;
    sm.item0 = 0;
    sm.serial = serial ? 1 : 0;
    int maxT = 1;
    if (!serial) {
        const int64_t T = std::min(nnp, std::min<int64_t>(n_task, 64));
        int q = 0;
        for (int64_t i = n; i > 0;) {                     // creation order: last body first, geopotential before ring
            --i;
            for (int kind = 0; kind < 2; ++kind) {
                const int32_t model = kind == 0 ? plan.np_gp[i] : plan.np_ring[i];
                if (model < 0) continue;
                int64_t c = 0;
                while ((c + 1) * nnp / T <= q) ++c;
                plan.items[q] = NPItem{(int32_t)i, kind, model, (int32_t)c, 0, (int32_t)(q * n)};
                ++q;
            }
        }
        sm.nitem = (int32_t)nnp;
        sm.T = (int32_t)T;
        maxT = std::max<int>(1, (int)T);
        if (nnp)
            for (int b = 0; b < n_main; ++b) plan.blk_recv[n_recv++] = plan.blk_main[b];
    }
    HIP_CHECK(hipMemcpy(base + kOffPlan, &plan, sizeof plan, hipMemcpyHostToDevice));

    const NaPlan *dp = (const NaPlan *)(base + kOffPlan);
    View v;
    v.F = (double *)(base + kOffF);
    v.nbp = (n + 63) / 64 * 64;
    v.nb = (int)n;
    v.sys = &dp->sys;
    v.items = dp->items;
    v.gps = gps;
    v.rings = rings;
    v.np_gp = dp->np_gp;
    v.np_ring = dp->np_ring;
    v.bsys = dp->bsys;
    v.perm = nullptr;                                     // (read only by the four-lane and one-thread-per-body deform kernels)
    v.srow_gp = dp->srow_gp;
    v.srow_ring = dp->srow_ring;
    v.san = n_sitems ? (const double *)(base + kOffAn) : nullptr;      // rows of at most 64 models x n partners
    v.an = (double *)(base + kOffAn);
    v.part = (double *)(base + kOffPart);
    v.t0 = &dp->t0;
    const hipStream_t st = 0;
    hipLaunchKernelGGL(k_na_gather, dim3((unsigned)((v.nbp + 63) / 64)), dim3(64), 0, st, masses, (int)n, v.nbp, v.F);
    // the kernels of the resident path for a batch of one system: one
    // workgroup per body (the item fields in extra workgroups of the deform
    // launch), then the main force in the order asked for and the non-point pass
    const int rows = serial ? 0 : (int)nnp;
    launch_deform_wpb(v, dp->work, (int)n, (int)n, 1, rows, 0, st);
    if (serial) {
        if (n_sitems) {
            View vs = v;
            vs.items = dp->sitems;
            vs.an = (double *)(base + kOffAn);
            launch_np_field(vs, n_sitems, (int)n, st);
        }
        launch_main_serial(v, dp->blk_main, n_main, st);
    } else {
        launch_main_wpb(v, dp->work, (int)n, (int)n, false, 0, st);
        if (nnp) launch_np(v, (int)nnp, (int)n, maxT, dp->blk_recv, n_recv, false, st);
    }
    hipLaunchKernelGGL(k_na_scatter, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, v.F, v.nbp, (int)n, masses, out);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
    return EPH_OK;
}
