// gfx950 kernels of the force evaluation and the RK12 stage engine.
//
// One thread owns one body i and walks its partners j in increasing index
// order through LDS-staged tiles (every lane reads the same tile entry, a
// broadcast), so each per-body sum is accumulated in exactly the order of the
// reference CPU path (src/physics/mass.cpp:112-157, 309-374) and the result
// is bit-identical to it.  Work is FP64 on the vector ALUs; there is no
// dense contraction, so no MFMA.
#include "eph_kernels.hpp"

#include <hip/hip_runtime.h>

namespace eph {

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------
__constant__ double2 c_rk[325];                    // Feagin RK12, {hi, lo}
// Non-zero coefficients of each row k = 1..25: entries c_row0[k] ..
// c_row0[k + 1] - 1 of c_nzj hold their column j, ascending (set with c_rk)
__constant__ int c_nzj[325];
__constant__ int c_row0[27];

__device__ const float g_fz[32 * 64 * 16] = {
#include "../tables/ring_fzcorr.inc"
};
#include "../tables/ring_pade.inc"
__constant__ double c_xserkk[][2] = EPH_PADE_XSERKK;
__constant__ double c_xserkm[][2] = EPH_PADE_XSERKM;
__constant__ double c_xser0[][2] = EPH_PADE_XSER0;
__constant__ double c_kserkk[][2] = EPH_PADE_KSERKK;
__constant__ double c_kserkm[][2] = EPH_PADE_KSERKM;
__constant__ double c_kser0[][2] = EPH_PADE_KSER0;

struct DevSin {
    __device__ double operator()(double x) const { return ::sin(x); }
};

// ---------------------------------------------------------------------------
// field access
// ---------------------------------------------------------------------------
struct Acc {
    double *F;
    int64_t P;
    __device__ double ld(int f, int g) const { return F[(int64_t)f * P + g]; }
    __device__ void st(int f, int g, double v) const { F[(int64_t)f * P + g] = v; }
    __device__ v3 ldv(int f, int g) const { return v3{ld(f, g), ld(f + 1, g), ld(f + 2, g)}; }
    __device__ void stv(int f, int g, v3 a) const { st(f, g, a.x); st(f + 1, g, a.y); st(f + 2, g, a.z); }
    __device__ m3 ldm(int f, int g) const { return m3{ldv(f, g), ldv(f + 3, g), ldv(f + 6, g)}; }
    __device__ void stm(int f, int g, const m3 &a) const { stv(f, g, a.x); stv(f + 3, g, a.y); stv(f + 6, g, a.z); }
    __device__ dd ldd(int f, int g) const { return dd{ld(f, g), ld(f + 1, g)}; }
    __device__ void std_(int f, int g, dd a) const { st(f, g, a.hi); st(f + 1, g, a.lo); }
    __device__ ddv3 lddv(int f, int g) const { return ddv3{ldd(f, g), ldd(f + 2, g), ldd(f + 4, g)}; }
    __device__ void stddv(int f, int g, ddv3 a) const { std_(f, g, a.x); std_(f + 2, g, a.y); std_(f + 4, g, a.z); }
    __device__ ddm3 lddm(int f, int g) const { return ddm3{lddv(f, g), lddv(f + 6, g), lddv(f + 12, g)}; }
    __device__ void stddm(int f, int g, const ddm3 &a) const { stddv(f, g, a.x); stddv(f + 6, g, a.y); stddv(f + 12, g, a.z); }
};

// ---------------------------------------------------------------------------
// ring::sum (src/physics/ring.impl:2070-2196) on the device tables
// ---------------------------------------------------------------------------
struct DevRingTab {
    __device__ static float fz(float x, float y, int kh) { return fzcorr_tab(g_fz, x, y, kh); }
    __device__ static double xserkk(double x) { return padesum(x, c_xserkk); }
    __device__ static double xserkm(double x) { return padesum(x, c_xserkm); }
    __device__ static double xser0(double x) { return padesum(x, c_xser0); }
    __device__ static double kserkk(double x) { return padesum(x, c_kserkk); }
    __device__ static double kserkm(double x) { return padesum(x, c_kserkm); }
    __device__ static double kser0(double x) { return padesum(x, c_kser0); }
};
__device__ v3 ring_sum(const eph_ring &rg, double host_R, v3 r) { return ring_sum_t<DevRingTab>(rg, host_R, r); }

// world-frame acceleration (per unit GM of the non-point body p) at offset r
// from p: geopotential (mass.cpp:186-188) or ring (mass.cpp:200-203)
__device__ __forceinline__ v3 nonpoint_field_inl(const View &d, const Acc &A, int p, int kind, int model, v3 r) {
    double R = A.ld(F_RAD, p);
    if (kind == 0) {
        m3 fmis = todbl(A.lddm(F_S, p));
        v3 lr = tolocal(fmis, r);
        return fmis * gp_sum(d.gps[model], R, lr);
    }
    v3 migl = todbl(A.lddv(F_GL, p));
    v3 x = asc_node(migl), z = unit(migl);
    m3 fgls{x, cross(z, x), z};
    v3 lr = tolocal(fgls, r);
    return fgls * ring_sum(d.rings[model], R, lr);
}
// The same out of line, for the kernels that already need every register
// (serial order, small systems): they pay a call instead.
__device__ __noinline__ v3 nonpoint_field_call(const View &d, const Acc &A, int p, int kind, int model, v3 r) {
    return nonpoint_field_inl(d, A, p, kind, model, r);
}


// FINALIZE_RELATIVITY and RING_CORRECTION (mass.impl:133-152) of body g with
// its final daccel/dtorque; writes phi, naccel, daccel, dtorque
__device__ void finalize_with(const View &d, const Acc &A, int g, v3 ga, v3 da, v3 dq) {
    A.st(F_PHI, g, A.ld(F_PHIC, g) * kC2);
    const v3 nacc = A.ldv(F_NACCC, g) * kC2;
    A.stv(F_NACC, g, nacc);
    const int rg = d.np_ring[g];
    if (rg >= 0) ring_correction(d.rings[rg], A.ld(F_R2, g), todbl(A.lddv(F_GL, g)), ga, nacc, da, dq);
    A.stv(F_DACC, g, da);
    A.stv(F_DTQ, g, dq);
}
__device__ void finalize_body(const View &d, const Acc &A, int g, v3 da, v3 dq) {
    finalize_with(d, A, g, A.ldv(F_GACC, g), da, dq);
}

// Field of non-point item `row` at partner jl of its system, per unit GM of
// the item's body (k_np_field; also run by extra workgroups of the deform
// launches, since it needs only the stage's positions and frames).
__device__ __forceinline__ void np_field_one(const View &d, const Acc &A, int row, int jl) {
    const NPItem it = d.items[row];
    const SysMeta sm = d.sys[it.sys];
    if (jl >= sm.n) return;
    const int gj = sm.off + jl;
    v3 an{0, 0, 0};
    if (gj != it.body) {
        const v3 r = diff(A.lddv(F_R, gj), A.lddv(F_R, it.body));
        an = nonpoint_field_inl(d, A, it.body, it.kind, it.model, r);
    }
    double *o = d.an + 3 * ((int64_t)it.an_off + jl);
    o[0] = an.x; o[1] = an.y; o[2] = an.z;
}

// Extra workgroups appended to a deform launch: each wavefront past the
// deform workgroups covers 64 partners of one item row.  Units are taken
// chunk-major, so the chunks of one row (a ring's are costly) run in
// different workgroups and on different CUs.
struct NpTail { int ndef, nrows, chunks; };           // chunks of 64 partners per row
__device__ __forceinline__ bool np_tail(const View &d, const NpTail &np) {
    if ((int)blockIdx.x < np.ndef) return false;
    if (np.nrows <= 0) return true;
    const int u = ((int)blockIdx.x - np.ndef) * ((int)blockDim.x >> 6) + ((int)threadIdx.x >> 6);
    const int c = u / np.nrows, row = u - c * np.nrows;
    if (c < np.chunks) np_field_one(d, Acc{d.F, d.nbp}, row, c * 64 + ((int)threadIdx.x & 63));
    return true;
}

// A round's partner block, loaded by the whole wavefront one round ahead:
// slot s = lane + 64 m holds field s / L of partner q0 + s % L (= lane % L),
// so the loads of one field are L consecutive doubles, and the global-memory
// latency of the next round overlaps this round's arithmetic.  In LDS,
// partner u's fields start at u * P (P odd: the partners' rows start in
// different banks).
template <int L, int NF>
struct PartnerBlock {
    static constexpr int NS = (L * NF + 63) / 64;
    static constexpr int P = NF | 1;
    int64_t fo[NS];
    int dst[NS];
    bool has[NS];
    double pf[NS];
    __device__ __forceinline__ void init(const int *fields, int64_t nbp, int lane) {
#pragma unroll
        for (int m = 0; m < NS; ++m) {
            const int sl = lane + 64 * m, f = sl / L;
            has[m] = sl < L * NF;
            fo[m] = has[m] ? (int64_t)fields[f] * nbp : 0;
            dst[m] = (sl % L) * P + f;
        }
    }
    __device__ __forceinline__ void fetch(const double *__restrict__ F, int off, int n, int q0, int k) {
        const int jl = q0 + k;
#pragma unroll
        for (int m = 0; m < NS; ++m) pf[m] = (has[m] && jl < n) ? F[fo[m] + off + jl] : 0.0;
    }
    __device__ __forceinline__ void store(double *jb) const {
#pragma unroll
        for (int m = 0; m < NS; ++m)
            if (has[m]) jb[dst[m]] = pf[m];
    }
};

// ---------------------------------------------------------------------------
// deform + spin fixed point + PREPARE_RELATIVITY (mass.cpp:78-135)
// ---------------------------------------------------------------------------
struct JDeform { double rxh, rxl, ryh, ryl, rzh, rzl, vxh, vxl, vyh, vyl, vzh, vzl, GM, sR2, tdf, pad; };

// Per-body constants of UPDATE_HARMONICS (mass.impl:48-61) from the state
__device__ __forceinline__ Harmonics harmonics_load(const Acc &Ac, int g) {
    Harmonics h;
    h.set(Ac.ld(F_K2R, g), Ac.ld(F_RAD, g), Ac.ld(F_R2, g), Ac.ld(F_A, g), Ac.ld(F_GM, g), Ac.ld(F_EXJ2, g),
          Ac.ldm(F_CST, g), todbl(Ac.lddm(F_S, g)));
    return h;
}

// UPDATE_ANGULAR_VELOCITY (mass.impl:63-72); returns should_break
__device__ __forceinline__ bool update_spin(const Acc &A, int g, const m3 &GI, v3 &w, const ddv3 &GL) {
    bool gl_from_w;
    const bool brk = spin_step(GI, w, GL, gl_from_w);
    if (gl_from_w) A.stddv(F_GL, g, mkddv(GI * w));
    return brk;
}

__device__ __forceinline__ void load_jdeform(const Acc &A, JDeform &T, int gj) {
    T.rxh = A.ld(F_R + 0, gj); T.rxl = A.ld(F_R + 1, gj);
    T.ryh = A.ld(F_R + 2, gj); T.ryl = A.ld(F_R + 3, gj);
    T.rzh = A.ld(F_R + 4, gj); T.rzl = A.ld(F_R + 5, gj);
    T.vxh = A.ld(F_V + 0, gj); T.vxl = A.ld(F_V + 1, gj);
    T.vyh = A.ld(F_V + 2, gj); T.vyl = A.ld(F_V + 3, gj);
    T.vzh = A.ld(F_V + 4, gj); T.vzl = A.ld(F_V + 5, gj);
    T.GM = A.ld(F_GM, gj); T.sR2 = A.ld(F_SR2, gj); T.tdf = A.ld(F_TDF, gj);
}

__device__ __forceinline__ void deform_write(const Acc &A, int g, int prepare, const ddv3 &vi, double phi, v3 nacc,
                                             const m3 &C, const m3 &GI, v3 w) {
    A.stm(F_CPOT, g, C);
    A.stm(F_GI, g, GI);
    A.stv(F_W, g, w);
    if (prepare) {                                         // PREPARE_RELATIVITY, mass.impl:76-80
        const v3 beta = todbl(vi) / kC;
        A.stv(F_BETA, g, beta);
        A.st(F_BETA2, g, dot(beta, beta));
        A.st(F_PHIC, g, phi / kC2);
        A.stv(F_NACCC, g, nacc / kC2);
    } else {                                               // plain msystem::deform()
        A.st(F_PHI, g, phi);
        A.stv(F_NACC, g, nacc);
    }
}

// Bodies without delayed tides (k2 == 0, or tide_delay == 0): the partner sum
// of deform_by does not depend on the spin, so the up-to-four passes of the
// spin fixed point (mass.cpp:123-129) all re-add the same terms; it is summed
// once and only UPDATE_HARMONICS / UPDATE_ANGULAR_VELOCITY are iterated.
// blk[b] = {system, first slot, slot count}; slots map to bodies via d.perm.
template <int BS>
__global__ __launch_bounds__(BS) void k_deform_light(View d, const int4 *__restrict__ blk, int prepare) {
    const Acc A{d.F, d.nbp};
    const int4 bk = blk[blockIdx.x];
    const SysMeta sm = d.sys[bk.x];
    const int n = sm.n, off = sm.off;
    const bool valid = (int)threadIdx.x < bk.z;
    const int gi = valid ? d.perm[off + bk.y + threadIdx.x] : off;
    const int il = gi - off;
    __shared__ JDeform tile[BS];
    const ddv3 ri = A.lddv(F_R, gi);
    const double sR2 = A.ld(F_SR2, gi), k2 = A.ld(F_K2, gi);
    const bool ndelay = k2 != 0;
    double phi = 0;
    v3 nacc{0, 0, 0};
    m3 Csum = mzero();
    for (int t0 = 0; t0 < n; t0 += BS) {
        __syncthreads();
        if (t0 + (int)threadIdx.x < n) load_jdeform(A, tile[threadIdx.x], off + t0 + (int)threadIdx.x);
        __syncthreads();
        if (!valid) continue;
        const int tn = min(BS, n - t0);
        for (int q = 0; q < tn; ++q) {
            if (t0 + q == il) continue;
            const JDeform &T = tile[q];
            const ddv3 rj{dd{T.rxh, T.rxl}, dd{T.ryh, T.ryl}, dd{T.rzh, T.rzl}};
            const v3 r = diff(rj, ri);
            const double rr2 = 1 / smax(dot(r, r), sR2 + T.sR2);
            const double rr = sqrt(rr2);
            const double tp_dphi = rr * T.GM;
            const double tp_dg = rr2 * tp_dphi;
            phi -= tp_dphi;                                // _NK2 (mass.impl:38-46) / _NDELAY (:25-36)
            nacc += tp_dg * r;
            if (ndelay) {
                const double fmj = -3 * tp_dg;
                Csum = Csum - (fmj * k2) * subdiag(outer(rr2 * r, r), 1.0 / 3);
            }
        }
    }
    if (!valid) return;
    Harmonics H;
    H = harmonics_load(A, gi);
    v3 w = A.ldv(F_W, gi);
    const ddv3 GL = A.lddv(F_GL, gi);
    m3 C, GI;
    int iters = 4;                                         // MAX_ANGULAR_VELOCITY_ITER
    for (;;) {
        H.apply(Csum, w, C, GI);
        if (!prepare || update_spin(A, gi, GI, w, GL) || --iters == 0) break;
    }
    deform_write(A, gi, prepare, A.lddv(F_V, gi), phi, nacc, C, GI, w);
}

// The same for large batches with four lanes per body (16 bodies per
// wavefront): lane k evaluates partner q0 + k of every group of four and
// leaves its terms in LDS; lane c then adds the group's terms, partner by
// partner, to chains c, c + 4, c + 8 (and lane 0 to chain 12), where the
// chains are phi (0), naccel (1-3) and the tidal matrix (4-12), the
// subtracted terms stored negated.  The self pair and partners past the end
// add +0, an exact identity.
constexpr int kDefK4Pitch = 13;
// fields of a deform partner: r (hi, lo), sR2, GM
constexpr int kJDefF = 8;
__constant__ int c_jdef_field[kJDefF] = {F_R + 0, F_R + 1, F_R + 2, F_R + 3, F_R + 4, F_R + 5, F_SR2, F_GM};

// Block `bid` of the list; term holds 64 x kDefK4Pitch doubles.  Its launch
// runs three wavefronts per SIMD (168 registers; the few values the compiler
// then keeps in scratch are stored before the partner loop and reloaded
// after it).
__device__ __forceinline__ void deform_k4_body(View d, const int4 *__restrict__ blk, int prepare, int bid, double *term) {
    const Acc A{d.F, d.nbp};
    const int4 bk = blk[bid];
    if (bk.z <= 0) return;                                 // padding of an XCD-dealt list
    const SysMeta sm = d.sys[bk.x];
    const int n = sm.n, off = sm.off;
    const int lane = (int)threadIdx.x, b = lane >> 2, k = lane & 3;
    const bool valid = b < bk.z;
    const int gi = valid ? d.perm[off + bk.y + b] : off;
    const int il = gi - off;
    using Blk = PartnerBlock<4, kJDefF>;
    __shared__ double jb[4 * Blk::P];
    Blk pb;
    pb.init(c_jdef_field, d.nbp, lane);
    pb.fetch(d.F, off, n, 0, k);
    pb.store(jb);
    const ddv3 ri = A.lddv(F_R, gi);
    const double sR2 = A.ld(F_SR2, gi), k2 = A.ld(F_K2, gi);
    const bool ndelay = valid && k2 != 0;
    const bool any_nd = __ballot(ndelay) != 0;             // wave-uniform
    double acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
    const double *grp = term + (b * 4) * kDefK4Pitch;
    __syncthreads();
    for (int q0 = 0; q0 < n; q0 += 4) {
        const int jl = q0 + k;
        const bool more = q0 + 4 < n;
        if (more) pb.fetch(d.F, off, n, q0 + 4, k);
        double t[13];
#pragma unroll
        for (int r = 0; r < 13; ++r) t[r] = 0.0;
        if (valid && jl < n && jl != il) {
            const double *J = jb + k * Blk::P;
            const ddv3 rj{dd{J[0], J[1]}, dd{J[2], J[3]}, dd{J[4], J[5]}};
            const v3 r = diff(rj, ri);
            const double rr2 = 1 / smax(dot(r, r), sR2 + J[6]);
            const double rr = sqrt(rr2);
            const double tp_dphi = rr * J[7];
            const double tp_dg = rr2 * tp_dphi;
            const v3 na = tp_dg * r;
            t[0] = -tp_dphi;                               // _NK2 (mass.impl:38-46) / _NDELAY (:25-36)
            t[1] = na.x; t[2] = na.y; t[3] = na.z;
            if (ndelay) {
                const double fmj = -3 * tp_dg;
                const m3 tm = (fmj * k2) * subdiag(outer(rr2 * r, r), 1.0 / 3);
                const double *tp = &tm.x.x;
#pragma unroll
                for (int e = 0; e < 9; ++e) t[4 + e] = -tp[e];
            }
        }
        if (any_nd) {
#pragma unroll
            for (int r = 0; r < 13; ++r) term[lane * kDefK4Pitch + r] = t[r];
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) term[lane * kDefK4Pitch + r] = t[r];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double *tr = grp + u * kDefK4Pitch;
            acc0 += tr[k];
            if (any_nd) {
                acc1 += tr[4 + k];
                acc2 += tr[8 + k];
                if (k == 0) acc3 += tr[12];
            }
        }
        if (more) pb.store(jb);                            // every lane has read this round's block
        __syncthreads();
    }
    // gather the body's chains into its lane 0
    const int base = lane & ~3;
    double v[3][3];
#pragma unroll
    for (int o = 1; o <= 3; ++o) {
        v[0][o - 1] = __shfl(acc0, base + o, 64);
        v[1][o - 1] = __shfl(acc1, base + o, 64);
        v[2][o - 1] = __shfl(acc2, base + o, 64);
    }
    if (!valid || k != 0) return;
    const double phi = acc0;
    const v3 nacc{v[0][0], v[0][1], v[0][2]};
    const m3 Csum{v3{acc1, v[1][0], v[1][1]}, v3{v[1][2], acc2, v[2][0]}, v3{v[2][1], v[2][2], acc3}};
    Harmonics H;
    H = harmonics_load(A, gi);
    v3 w = A.ldv(F_W, gi);
    const ddv3 GL = A.lddv(F_GL, gi);
    m3 C, GI;
    int iters = 4;                                         // MAX_ANGULAR_VELOCITY_ITER
    for (;;) {
        H.apply(Csum, w, C, GI);
        if (!prepare || update_spin(A, gi, GI, w, GL) || --iters == 0) break;
    }
    deform_write(A, gi, prepare, A.lddv(F_V, gi), phi, nacc, C, GI, w);
}

// Bodies with constant-time-lag tides (k2 != 0, tide_delay != 0): the tidal
// phase depends on the spin, so each fixed-point pass re-sums the partners
// (DAMPED_TIDAL_DEFORMATION_MATRIX, mass.impl:1-23).
template <int BS>
__global__ __launch_bounds__(BS) void k_deform_damped(View d, const int4 *__restrict__ blk, int prepare) {
    const Acc A{d.F, d.nbp};
    const int4 bk = blk[blockIdx.x];
    const SysMeta sm = d.sys[bk.x];
    const int n = sm.n, off = sm.off;
    const bool valid = (int)threadIdx.x < bk.z;
    const int gi = valid ? d.perm[off + bk.y + threadIdx.x] : off;
    const int il = gi - off;
    __shared__ JDeform tile[BS];
    const ddv3 ri = A.lddv(F_R, gi), vi = A.lddv(F_V, gi);
    const double sR2 = A.ld(F_SR2, gi), k2 = A.ld(F_K2, gi), td = A.ld(F_TD, gi);
    const ddv3 GL = A.lddv(F_GL, gi);
    Harmonics H;
    H = harmonics_load(A, gi);
    v3 w = A.ldv(F_W, gi);
    double phi = 0;
    v3 nacc{0, 0, 0};
    m3 C = mzero(), GI = mzero();
    int iters_left = 4;
    bool active = valid;
    for (;;) {
        m3 Csum = mzero();
        if (active) { phi = 0; nacc = v3{0, 0, 0}; }
        for (int t0 = 0; t0 < n; t0 += BS) {
            __syncthreads();
            if (t0 + (int)threadIdx.x < n) load_jdeform(A, tile[threadIdx.x], off + t0 + (int)threadIdx.x);
            __syncthreads();
            if (!active) continue;
            const int tn = min(BS, n - t0);
            for (int q = 0; q < tn; ++q) {
                if (t0 + q == il) continue;
                const JDeform &T = tile[q];
                const ddv3 rj{dd{T.rxh, T.rxl}, dd{T.ryh, T.ryl}, dd{T.rzh, T.rzl}};
                const ddv3 vj{dd{T.vxh, T.vxl}, dd{T.vyh, T.vyl}, dd{T.vzh, T.vzl}};
                const v3 r = diff(rj, ri);
                const v3 v = diff(vj, vi);
                const double rr2 = 1 / smax(dot(r, r), sR2 + T.sR2);
                const double rr = sqrt(rr2);
                const double tp_dphi = rr * T.GM;
                phi -= tp_dphi;
                const double tp_dg = rr2 * tp_dphi;
                nacc += tp_dg * r;
                const double fmj = -3 * tp_dg;
                v3 dw = rr2 * cross(r, cross(w, r) - v);
                double dwt = normalize(dw);
                dwt *= td * T.tdf;
                const double ecwt2 = 1 + 4 * dwt * dwt,
                             secwt2 = sqrt(ecwt2),
                             x0 = sqrt((1 + secwt2) / (2 * ecwt2)),
                             y0 = dwt * sqrt(2 / (ecwt2 * (1 + secwt2))),
                             z02 = dwt * dwt * (2 / (ecwt2 + secwt2));
                const v3 dr = x0 * r + y0 * cross(dw, r);
                m3 t = outer(rr2 * dr, dr) - outer(z02 * dw, dw);
                t = adddiag(t, z02 - 1.0 / 3);
                Csum = Csum - (fmj * k2) * t;
            }
        }
        bool again = false;
        if (active) {
            H.apply(Csum, w, C, GI);
            const bool brk = !prepare || update_spin(A, gi, GI, w, GL);
            again = !brk && --iters_left > 0;
            active = again;
        }
        if (!__syncthreads_or(again)) break;
    }
    if (!valid) return;
    deform_write(A, gi, prepare, vi, phi, nacc, C, GI, w);
}

// ---------------------------------------------------------------------------
// main force, thread-pool summation order (mass.cpp:136-157)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_jmain(const Acc &A, JMain &T, int gj) {
    T.rxh = A.ld(F_R + 0, gj); T.rxl = A.ld(F_R + 1, gj);
    T.ryh = A.ld(F_R + 2, gj); T.ryl = A.ld(F_R + 3, gj);
    T.rzh = A.ld(F_R + 4, gj); T.rzl = A.ld(F_R + 5, gj);
    T.GM = A.ld(F_GM, gj); T.sR2 = A.ld(F_SR2, gj);
    T.phi = A.ld(F_PHIC, gj); T.beta2 = A.ld(F_BETA2, gj);
    T.bx = A.ld(F_BETA, gj); T.by = A.ld(F_BETA + 1, gj); T.bz = A.ld(F_BETA + 2, gj);
    T.nx = A.ld(F_NACCC, gj); T.ny = A.ld(F_NACCC + 1, gj); T.nz = A.ld(F_NACCC + 2, gj);
#pragma unroll
    for (int q = 0; q < 9; ++q) T.C[q] = A.ld(F_CPOT + q, gj);
    T.R2 = A.ld(F_R2, gj);
    T.GLx = todbl(dd{A.ld(F_GL + 0, gj), A.ld(F_GL + 1, gj)});
    T.GLy = todbl(dd{A.ld(F_GL + 2, gj), A.ld(F_GL + 3, gj)});
    T.GLz = todbl(dd{A.ld(F_GL + 4, gj), A.ld(F_GL + 5, gj)});
    T.lum = A.ld(F_LUM, gj);
}

__device__ __forceinline__ IMain load_imain(const Acc &A, int gi) {
    IMain I;
    I.r = A.lddv(F_R, gi);
    I.sR2 = A.ld(F_SR2, gi); I.phi = A.ld(F_PHIC, gi); I.beta2 = A.ld(F_BETA2, gi);
    I.R2 = A.ld(F_R2, gi); I.GM = A.ld(F_GM, gi); I.rR2G4c = A.ld(F_RR2G4C, gi);
    I.beta = A.ldv(F_BETA, gi);
    I.GL = todbl(A.lddv(F_GL, gi));
    I.C = A.ldm(F_CPOT, gi);
    return I;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_main_pool(View d, const int2 *__restrict__ blk) {
    const Acc A{d.F, d.nbp};
    const int s = blk[blockIdx.x].x, i0 = blk[blockIdx.x].y;
    const SysMeta sm = d.sys[s];
    const int n = sm.n, off = sm.off;
    const int il = i0 + (int)threadIdx.x;
    const bool valid = il < n;
    const int gi = off + (valid ? il : 0);
    __shared__ JMain tile[BS];
    const IMain I = load_imain(A, gi);
    MainAcc acc;
    acc.mind = A.ld(F_MIND, gi);
    acc.maxi = A.ld(F_MAXI, gi);
    for (int t0 = 0; t0 < n; t0 += BS) {
        __syncthreads();
        if (t0 + (int)threadIdx.x < n) load_jmain(A, tile[threadIdx.x], off + t0 + (int)threadIdx.x);
        __syncthreads();
        if (!valid) continue;
        const int tn = min(BS, n - t0);
        for (int q = 0; q < tn; ++q) {
            if (t0 + q == il) continue;
            acc.add(pair_terms(I, tile[q]));
        }
    }
    if (!valid) return;
    A.stv(F_GACC, gi, acc.ga);
    A.stv(F_DACC, gi, acc.da);
    A.stv(F_DTQ, gi, acc.dq);
    A.st(F_MIND, gi, acc.mind);
    A.st(F_MAXI, gi, acc.maxi);
    if (sm.nitem == 0) {                                   // FINALIZE_RELATIVITY
        A.st(F_PHI, gi, I.phi * kC2);
        A.stv(F_NACC, gi, A.ldv(F_NACCC, gi) * kC2);
    }
}


// ---------------------------------------------------------------------------
// throughput path for large batches: four lanes per body, 16 bodies per
// wavefront.  Lane k of a body evaluates partner q0 + k of every group of
// four and leaves the pair's terms in LDS; lanes 0-2 then add the group's
// terms, partner by partner, to the x, y or z component of daccel (own and
// partner's quadrupole, -LT own spin, +LT partner's spin, radiation), gaccel
// and dtorque (r x dg, -LT torque) -- the reference order of every chain.
// The self pair and partners past the end add +0, an exact identity (the
// sums start at +0 and never hold -0); the running maxima are order-free.
// Four times as many wavefronts as one thread per body keep every SIMD busy
// and balanced across systems of different sizes.
// ---------------------------------------------------------------------------
constexpr int kK4Pitch = 25;

// pair_terms for the four-lane kernel, which only adds the terms to sums that
// start at +0: a term that is exactly zero for finite inputs -- the own or
// the partner's quadrupole when there is none, the partner-spin LT term of a
// partner without spin, the radiation of a dark partner -- is left at +0
// without being formed (adding a zero of either sign to such a sum is an
// exact identity).  The branches are uniform whenever the wavefront's 16
// bodies or 4 partners agree, which they mostly do.
__device__ __forceinline__ PairTerms pair_terms_k4(const IMain &I, bool own_quad, const JMain &T) {
    PairTerms o;
    v3 unused{0, 0, 0};
    const PairGeo p = relativity(I, T, unused);
    o.G = p.G;
    o.rr = p.rr;
    o.tp_dg = p.tp_dg;
    o.dg = v3{0, 0, 0};
    o.rxdg = v3{0, 0, 0};
    if (own_quad) {
        o.dg = (T.GM * p.rr3 * p.rr2 * I.R2) * quad_field(I.C, p);
        o.rxdg = cross(p.r, o.dg);
    }
    o.gj = v3{0, 0, 0};
    bool cj = false;
#pragma unroll
    for (int q = 0; q < 9; ++q) cj = cj || T.C[q] != 0;
    if (cj && T.R2 != 0) {
        const m3 Cj{v3{T.C[0], T.C[1], T.C[2]}, v3{T.C[3], T.C[4], T.C[5]}, v3{T.C[6], T.C[7], T.C[8]}};
        o.gj = (T.GM * p.rr3 * p.rr2 * T.R2) * quad_field(Cj, p);
    }
    const double rcr32 = 2 * T.GM / kC * p.rr3;
    const v3 GLb = cross(I.GL, p.b);
    o.lt1 = rcr32 * (GLb - (3 * dot(GLb, p.r) * p.rr2) * p.r);
    o.lt2 = rcr32 * cross(I.GL, cross(p.r, p.b));
    const v3 GLj{T.GLx, T.GLy, T.GLz};
    o.lt3 = v3{0, 0, 0};
    if (GLj.x != 0 || GLj.y != 0 || GLj.z != 0) o.lt3 = cross(rcr32 * (GLj - (3 * dot(GLj, p.r) * p.rr2) * p.r), p.b);
    o.rad = v3{0, 0, 0};
    if (T.lum != 0) o.rad = (T.lum * I.rR2G4c * p.rr3 / I.GM) * p.r;
    return o;
}

// block `bid` of the list; term holds 64 x kK4Pitch doubles
__device__ __forceinline__ void main_k4_body(View d, const int2 *__restrict__ blk, int bid, double *term) {
    const Acc A{d.F, d.nbp};
    const int s = blk[bid].x, i0 = blk[bid].y;             // bodies i0 .. i0 + 15
    if (i0 < 0) return;                                    // padding of an XCD-dealt list
    const SysMeta sm = d.sys[s];
    const int n = sm.n, off = sm.off;
    const int lane = (int)threadIdx.x, b = lane >> 2, k = lane & 3;
    const int il = i0 + b;
    const bool valid = il < n;
    const int gi = off + (valid ? il : 0);
    const IMain I = load_imain(A, gi);
    bool own_quad = false;
    {
        const double *ci = &I.C.x.x;
#pragma unroll
        for (int q = 0; q < 9; ++q) own_quad = own_quad || ci[q] != 0;
        own_quad = own_quad && I.R2 != 0;
    }
    const int c = k < 3 ? k : 0;
    double a_d = 0, a_g = 0, a_q = 0;                      // daccel, gaccel, dtorque component c
    double mind = 0, maxi = 0;
    const double *grp = term + (b * 4) * kK4Pitch;
    for (int q0 = 0; q0 < n; q0 += 4) {
        const int jl = q0 + k;
        double t[24];
        if (valid && jl < n && jl != il) {
            JMain T;
            load_jmain(A, T, off + jl);
            const PairTerms pt = pair_terms_k4(I, own_quad, T);
            const double v[24] = {pt.G.x, pt.G.y, pt.G.z, pt.dg.x, pt.dg.y, pt.dg.z, pt.gj.x, pt.gj.y, pt.gj.z,
                                  -pt.lt1.x, -pt.lt1.y, -pt.lt1.z, pt.lt3.x, pt.lt3.y, pt.lt3.z,
                                  pt.rad.x, pt.rad.y, pt.rad.z, pt.rxdg.x, pt.rxdg.y, pt.rxdg.z,
                                  -pt.lt2.x, -pt.lt2.y, -pt.lt2.z};
#pragma unroll
            for (int r = 0; r < 24; ++r) t[r] = v[r];
            mind = checked_max_into(mind, pt.rr);
            maxi = checked_max_into(maxi, pt.tp_dg);
        } else {
#pragma unroll
            for (int r = 0; r < 24; ++r) t[r] = 0.0;
        }
#pragma unroll
        for (int r = 0; r < 24; ++r) term[lane * kK4Pitch + r] = t[r];
        __syncthreads();
        if (k < 3) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double *tr = grp + u * kK4Pitch;
                a_d += tr[3 + c];
                a_d += tr[6 + c];
                a_d += tr[9 + c];
                a_d += tr[12 + c];
                a_d += tr[15 + c];
                a_g += tr[c];
                a_q += tr[18 + c];
                a_q += tr[21 + c];
            }
        }
        __syncthreads();
    }
    mind = checked_max_into(mind, __shfl_xor(mind, 1, 64));
    mind = checked_max_into(mind, __shfl_xor(mind, 2, 64));
    maxi = checked_max_into(maxi, __shfl_xor(maxi, 1, 64));
    maxi = checked_max_into(maxi, __shfl_xor(maxi, 2, 64));
    if (!valid) return;
    if (k < 3) {
        A.st(F_GACC + c, gi, a_g);
        A.st(F_DACC + c, gi, a_d);
        A.st(F_DTQ + c, gi, a_q);
    } else {
        A.st(F_MIND, gi, checked_max_into(A.ld(F_MIND, gi), mind));
        A.st(F_MAXI, gi, checked_max_into(A.ld(F_MAXI, gi), maxi));
        if (sm.nitem == 0) {                               // FINALIZE_RELATIVITY
            A.st(F_PHI, gi, I.phi * kC2);
            A.stv(F_NACC, gi, A.ldv(F_NACCC, gi) * kC2);
        }
    }
}

__global__ __launch_bounds__(64) void k_main_k4(View d, const int2 *__restrict__ blk) {
    __shared__ double term[64 * kK4Pitch];
    main_k4_body(d, blk, (int)blockIdx.x, term);
}

// ---------------------------------------------------------------------------
// latency path: one workgroup of 1-4 wavefronts per body.  Each lane
// evaluates one partner of a 64-partner tile and leaves that pair's terms in
// LDS; then one lane per accumulator adds them in partner order, so every
// sum is formed exactly as in the in-order loop above.
//
// gaccel (+G per partner) is one chain per component over the tile's G rows.
// daccel and dtorque take up to five additions per partner, in the reference
// order of the pair (mass.cpp:147-155, RELATIVITY .. RADIATION_PRESSURE of
// mass.impl); slot k is one addition for each of the six lanes:
//   slot 0  +dg (own quadrupole)         +r x dg
//   slot 1  +gj (partner's quadrupole)    .
//   slot 2  -LT (own spin)               -LT torque
//   slot 3  +LT (partner's spin)          .
//   slot 4  +radiation                    .
// ('.' adds +0).  The lane that evaluates a partner stores only the slots
// with a non-zero term, as consecutive six-value entries: a sum that starts
// at +0 never holds -0 in round-to-nearest, so adding a zero of either sign
// is an exact identity and leaving it out changes nothing (most partners
// have no quadrupole, no spin or no luminosity, so the dependent chain that
// bounds this kernel is shorter).  The self pair and partners past the end
// store nothing (G: +0).  a - b is a + (-b) exactly, so subtracted terms are
// stored negated.  The running maxima do not depend on order (ties are equal
// values, NaN wins either way) and are reduced across the wave at the end.
// ---------------------------------------------------------------------------
constexpr int kWpbPitch = 65;                        // G row pitch: rows start in different LDS banks
constexpr int kWpbSlots = 5;
constexpr int kWpbEnt = 6;                           // doubles per entry
constexpr int kWpbEntMax = 64 * kWpbSlots + 16;      // entries of one tile, padded to a multiple of 16
constexpr int kWpbTile = 3 * kWpbPitch + kWpbEnt * kWpbEntMax;

// acc += row[0] + ... + row[63], one dependent addition at a time in index
// order; the reads of the next sixteen values are issued before this
// sixteen are added (the scheduling barriers keep the compiler from
// gathering all reads ahead of one wait), so the chain does not wait for LDS.
__device__ __forceinline__ void chain_add64(const double *row, double &acc) {
    double b0[16], b1[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) b0[i] = row[i];
#pragma unroll
    for (int q = 0; q < 64; q += 32) {
#pragma unroll
        for (int i = 0; i < 16; ++i) b1[i] = row[q + 16 + i];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 16; ++i) acc += b0[i];
        if (q + 32 < 64) {
#pragma unroll
            for (int i = 0; i < 16; ++i) b0[i] = row[q + 32 + i];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 16; ++i) acc += b1[i];
    }
}
// acc += p[0] + p[S] + ... + p[(cnt - 1) * S], one dependent addition at a
// time in index order (cnt a multiple of 16, uniform across the wave); the
// reads of the next sixteen are issued before this sixteen are added.
template <int S>
__device__ __forceinline__ void chain_add_strided(const double *p, int cnt, double &acc) {
    if (cnt <= 0) return;
    double b0[16], b1[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) b0[i] = p[i * S];
    for (int e = 0; e < cnt; e += 32) {
        if (e + 16 < cnt) {
#pragma unroll
            for (int i = 0; i < 16; ++i) b1[i] = p[(e + 16 + i) * S];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 16; ++i) acc += b0[i];
        if (e + 16 >= cnt) break;
        if (e + 32 < cnt) {
#pragma unroll
            for (int i = 0; i < 16; ++i) b0[i] = p[(e + 32 + i) * S];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 16; ++i) acc += b1[i];
    }
}

__device__ __forceinline__ double wave_checked_max(double x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x = checked_max_into(x, __shfl_xor(x, m, 64));
    return x;
}

constexpr int kNpFusedMaxN = 512;     // systems up to this size take the fused non-point pass
constexpr int kNpFusedMaxItems = 64;

// Non-point pass of body i inside its own workgroup (mass.cpp:283-307),
// after k_np_field has left every item's field at every partner in d.an.
// Its inputs are gathered by spare wavefronts while wave 0 adds the last
// tile's chains, so that only the ordered additions follow the pair loop.
//
// Receiver side, lane q < nitem: GM of item q's body times its field at i
// (zero for i's own items, which the receiver sum skips).
__device__ __forceinline__ v3 np_recv_term(const View &d, const Acc &A, const SysMeta &sm, int il, int gi, int q,
                                           int &own_kind) {
    const NPItem it = d.items[sm.item0 + q];
    own_kind = it.body == gi ? it.kind : -1;
    if (it.body == gi) return v3{0, 0, 0};
    const double *a = d.an + 3 * ((int64_t)it.an_off + il);
    return A.ld(F_GM, it.body) * v3{a[0], a[1], a[2]};
}
// FINALIZE_RELATIVITY of a body with non-point items and the inputs of its
// RING_CORRECTION (mass.impl:133-152; finalize_with): phi and naccel are
// final already; ring: {GM_ratio, rmGL / (rmGL + GL_R2 * R2), mGL, naccel}
__device__ __forceinline__ void np_finalize_pre(const View &d, const Acc &A, int gi, double phic, double *ring8) {
    A.st(F_PHI, gi, phic * kC2);
    const v3 nacc = A.ldv(F_NACCC, gi) * kC2;
    A.stv(F_NACC, gi, nacc);
    const int rg = d.np_ring[gi];
    if (rg >= 0) {
        const eph_ring &mr = d.rings[rg];
        v3 mGL = todbl(A.lddv(F_GL, gi));
        const double rmGL = normalize(mGL);
        ring8[0] = mr.GM_ratio;
        ring8[1] = rmGL / (rmGL + mr.GL_R2 * A.ld(F_R2, gi));
        ring8[2] = mGL.x; ring8[3] = mGL.y; ring8[4] = mGL.z;
        ring8[5] = nacc.x; ring8[6] = nacc.y; ring8[7] = nacc.z;
    }
}
// LDS scratch of the non-point pass (in the pair-term tiles, free by then):
// rf, ring inputs, two own items' round sums, then per own item the
// products of every partner (6 doubles each)
constexpr int kNpHead = 15 * kNpFusedMaxItems + 8;
constexpr int kNpScratch = kNpHead + 6 * kNpFusedMaxN;

template <bool NP, int W>
__global__ __launch_bounds__(64 * W) void k_main_wpb(View d, const int2 *__restrict__ work) {
    static_assert(!NP || W * kWpbTile >= kNpScratch, "term tiles double as non-point scratch");
    const Acc A{d.F, d.nbp};
    const int s = work[blockIdx.x].x, il = work[blockIdx.x].y;
    const SysMeta sm = d.sys[s];
    const int n = sm.n, off = sm.off, gi = off + il;
    const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
    __shared__ double term[W * kWpbTile];
    __shared__ int count[W];
    __shared__ double wmax[2 * W];
    __shared__ double gsum[3];
    __shared__ int own_s[2];
    const IMain I = load_imain(A, gi);
    double mind0 = 0, maxi0 = 0;
    if (wv == 0) { mind0 = A.ld(F_MIND, gi); maxi0 = A.ld(F_MAXI, gi); }
    const int ring_i = NP ? d.np_ring[gi] : -1;            // read early: the tail needs it last
    double *myg = term + wv * kWpbTile, *mye = myg + 3 * kWpbPitch;
    double acc = 0, gacc = 0;                              // lanes 0-5 of wave 0: daccel, dtorque; gaccel
    double mind = 0, maxi = 0;                             // lane-local maxima (>= +0: identities)
    constexpr int kGWave = W > 1 ? 1 : 0;                  // the wave that adds the G chains
    // the waves that gather the non-point inputs during the last tile's chains
    constexpr int kHA = W > 2 ? 2 : W - 1, kHB = W > 3 ? 3 : kHA;
    const bool np = NP && sm.nitem > 0;
    v3 rfv{0, 0, 0}, tdA{0, 0, 0}, tqA{0, 0, 0}, tdB{0, 0, 0}, tqB{0, 0, 0};
    double ring8[8];
    int own0 = -1, own1 = -1;
    // with four wavefronts, waves 2 and 3 also form the first own item's
    // partner products during the last chains (partners lane + 128 m of their
    // 128 lanes; a second own item, a ring, is formed after the loop)
    constexpr bool kPre = W == 4;
    double P0[4][6];
    for (int t0 = 0; t0 < n; t0 += 64 * W) {
        const int jl = t0 + 64 * wv + lane;
        unsigned keep = 0;
        v3 e[2 * kWpbSlots];
        v3 G{0, 0, 0};
        if (jl < n && jl != il) {
            JMain T;
            load_jmain(A, T, off + jl);
            const PairTerms pt = pair_terms(I, T);
            G = pt.G;
            e[0] = pt.dg; e[1] = pt.rxdg;
            e[2] = pt.gj; e[3] = v3{0, 0, 0};
            e[4] = -pt.lt1; e[5] = -pt.lt2;
            e[6] = pt.lt3; e[7] = v3{0, 0, 0};
            e[8] = pt.rad; e[9] = v3{0, 0, 0};
            auto nz = [](const v3 &a) { return a.x != 0 || a.y != 0 || a.z != 0; };
#pragma unroll
            for (int k = 0; k < kWpbSlots; ++k)
                if (nz(e[2 * k]) || nz(e[2 * k + 1])) keep |= 1u << k;
            mind = checked_max_into(mind, pt.rr);
            maxi = checked_max_into(maxi, pt.tp_dg);
        }
        myg[lane] = G.x; myg[kWpbPitch + lane] = G.y; myg[2 * kWpbPitch + lane] = G.z;
        // this tile's entries: partner order, slot order within a partner
        int pos = 0, total = 0;
#pragma unroll
        for (int k = 0; k < kWpbSlots; ++k) {
            const uint64_t b = __ballot((keep >> k) & 1u);
            pos += __popcll(b & ((1ull << lane) - 1));
            total += __popcll(b);
        }
#pragma unroll
        for (int k = 0; k < kWpbSlots; ++k)
            if ((keep >> k) & 1u) {
                double *o = mye + kWpbEnt * pos++;
                o[0] = e[2 * k].x; o[1] = e[2 * k].y; o[2] = e[2 * k].z;
                o[3] = e[2 * k + 1].x; o[4] = e[2 * k + 1].y; o[5] = e[2 * k + 1].z;
            }
        const int padded = (total + 15) & ~15;
        if (lane < padded - total) {
            double *o = mye + kWpbEnt * (total + lane);
#pragma unroll
            for (int c = 0; c < kWpbEnt; ++c) o[c] = 0.0;
        }
        if (lane == 0) count[wv] = padded;
        __syncthreads();
        if (wv == 0 && lane < 6) {
            for (int u = 0; u < W; ++u)
                chain_add_strided<kWpbEnt>(term + u * kWpbTile + 3 * kWpbPitch + lane,
                                           __builtin_amdgcn_readfirstlane(count[u]), acc);
        }
        if (wv == kGWave && lane < 3) {
            const int nw = min(W, (n - t0 + 63) / 64);
            for (int u = 0; u < nw; ++u) chain_add64(term + u * kWpbTile + lane * kWpbPitch, gacc);
        }
        if (np && t0 + 64 * W >= n && (wv == kHA || (kPre && wv == kHB))) {
            int kind = -1;
            if (lane < sm.nitem) {
                const v3 f = np_recv_term(d, A, sm, il, gi, lane, kind);
                if (wv == kHA) rfv = f;
            }
            const uint64_t b0 = __ballot(kind == 0), b1 = __ballot(kind == 1);
            own0 = b0 ? __ffsll((unsigned long long)b0) - 1 : -1;
            own1 = b1 ? __ffsll((unsigned long long)b1) - 1 : -1;
            if (wv == kHA && lane == 0) np_finalize_pre(d, A, gi, I.phi, ring8);
            if (kPre && own0 >= 0) {
                const int off0 = d.items[sm.item0 + own0].an_off;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int jl = (wv - kHA) * 64 + lane + 128 * m;
#pragma unroll
                    for (int c = 0; c < 6; ++c) P0[m][c] = 0;
                    if (jl < n && jl != il) {
                        const int gj = off + jl;
                        const double GMj = A.ld(F_GM, gj);
                        const v3 rij = diff(A.lddv(F_R, gj), I.r);
                        const double *a = d.an + 3 * ((int64_t)off0 + jl);
                        const v3 an{a[0], a[1], a[2]};
                        const v3 ta = GMj * an, tt = GMj * cross(rij, an);
                        P0[m][0] = ta.x; P0[m][1] = ta.y; P0[m][2] = ta.z;
                        P0[m][3] = tt.x; P0[m][4] = tt.y; P0[m][5] = tt.z;
                    }
                }
            }
        }
        __syncthreads();
    }
    // the non-point inputs into the (now free) tiles
    double *rf = term, *ringS = rf + 3 * kNpFusedMaxItems, *partA = ringS + 8,
           *partB = partA + 6 * kNpFusedMaxItems, *pq = term + kNpHead;
    constexpr bool kBoth = W * kWpbTile >= kNpHead + 12 * kNpFusedMaxN;   // both own items' products at once
    static_assert(!kPre || kBoth, "the products formed during the chains use both items' regions");
    if (kPre && np && (wv == kHA || wv == kHB) && own0 >= 0) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int jl = (wv - kHA) * 64 + lane + 128 * m;
            if (jl < n) {
                double *o = pq + 6 * jl;
#pragma unroll
                for (int c = 0; c < 6; ++c) o[c] = P0[m][c];
            }
        }
    }
    if (np && wv == kHA) {
        if (lane < sm.nitem) { rf[3 * lane] = rfv.x; rf[3 * lane + 1] = rfv.y; rf[3 * lane + 2] = rfv.z; }
        if (lane == 0) {
            own_s[0] = own0; own_s[1] = own1;
#pragma unroll
            for (int k = 0; k < 8; ++k) ringS[k] = ring8[k];
        }
    }
    // gather the nine sums in lanes 0-8 of wave 0: gaccel, daccel, dtorque
    if (wv == kGWave && lane < 3) gsum[lane] = gacc;
    __syncthreads();
    if (wv == 0) {
        const double a6 = __shfl(acc, lane >= 3 ? lane - 3 : 0, 64);
        acc = lane < 3 ? gsum[lane] : a6;
    }
    mind = wave_checked_max(mind);
    maxi = wave_checked_max(maxi);
    if (lane == 0) { wmax[wv] = mind; wmax[W + wv] = maxi; }
    // own items (mass.cpp:298-307): the products GM_j * field and
    // GM_j * (r_j - r_i) x field of every partner (+0 for the self pair), then
    // round r < T subtracts those of partner slice (chain + r) mod T in order,
    // one lane per round and component
    if (np) {
        const int T = sm.T;
        const int tid = (int)threadIdx.x;
        const int q0 = own_s[0], q1 = own_s[1];
        for (int pass = 0; pass < (kBoth ? 1 : 2); ++pass) {
            const bool do0 = q0 >= 0 && (kBoth || pass == 0), do1 = q1 >= 0 && (kBoth || pass == 1);
            const bool f0 = do0 && !kPre;          // item 0's products are formed already (kPre)
            if (f0 || do1) {
                const int off0 = f0 ? d.items[sm.item0 + q0].an_off : 0;
                const int off1 = do1 ? d.items[sm.item0 + q1].an_off : 0;
                double *p0 = pq, *p1 = pq + (kBoth ? 6 * kNpFusedMaxN : 0);
                for (int jl = tid; jl < n; jl += 64 * W) {
                    v3 a0{0, 0, 0}, t0v{0, 0, 0}, a1{0, 0, 0}, t1v{0, 0, 0};
                    if (jl != il) {
                        const int gj = off + jl;
                        const double GMj = A.ld(F_GM, gj);
                        const v3 rij = diff(A.lddv(F_R, gj), I.r);
                        if (f0) {
                            const double *a = d.an + 3 * ((int64_t)off0 + jl);
                            const v3 an{a[0], a[1], a[2]};
                            a0 = GMj * an;
                            t0v = GMj * cross(rij, an);
                        }
                        if (do1) {
                            const double *a = d.an + 3 * ((int64_t)off1 + jl);
                            const v3 an{a[0], a[1], a[2]};
                            a1 = GMj * an;
                            t1v = GMj * cross(rij, an);
                        }
                    }
                    if (f0) {
                        double *o = p0 + 6 * jl;
                        o[0] = a0.x; o[1] = a0.y; o[2] = a0.z; o[3] = t0v.x; o[4] = t0v.y; o[5] = t0v.z;
                    }
                    if (do1) {
                        double *o = p1 + 6 * jl;
                        o[0] = a1.x; o[1] = a1.y; o[2] = a1.z; o[3] = t1v.x; o[4] = t1v.y; o[5] = t1v.z;
                    }
                }
            }
            __syncthreads();
            // item 0's rounds in wave 0, item 1's in wave 1 (both in wave 0 for one item at a time)
            const int k2 = kBoth ? tid >> 6 : pass;
            const int q = k2 == 0 ? q0 : q1;
            if ((kBoth ? k2 < 2 : tid < 64) && q >= 0 && (kBoth || (k2 == 0 ? do0 : do1))) {
                const double *pk = k2 == 0 ? pq : pq + (kBoth ? 6 * kNpFusedMaxN : 0);
                const int chain = d.items[sm.item0 + q].chain;
                double *part = k2 == 0 ? partA : partB;
                for (int u = tid & 63; u < 6 * T; u += 64) {
                    const int r = u / 6, c = u - 6 * r;
                    int kk = chain + r;
                    if (kk >= T) kk -= T;
                    const int j0 = kk * n / T, j1 = (kk + 1) * n / T;
                    double x = 0;
#pragma unroll 8
                    for (int jl = j0; jl < j1; ++jl) x -= pk[6 * jl + c];   // the self pair holds +0: an exact identity
                    part[6 * r + c] = x;
                }
            }
            if (!kBoth && pass == 0) __syncthreads();      // the products of item 1 overwrite item 0's
        }
    }
    __syncthreads();
    if (wv != 0) return;
    for (int u = 1; u < W; ++u) {
        mind = checked_max_into(mind, wmax[u]);
        maxi = checked_max_into(maxi, wmax[W + u]);
    }
    mind = checked_max_into(mind0, mind);
    maxi = checked_max_into(maxi0, maxi);
    if (lane < 3) A.st(F_GACC + lane, gi, acc);
    else if (lane == 9) A.st(F_MIND, gi, mind);
    else if (lane == 10) A.st(F_MAXI, gi, maxi);
    if (np) {
        const int nit = sm.nitem, T = sm.T;
        const int o0 = own_s[0], o1 = own_s[1];
        // (a) receiver side (mass.cpp:283-297): rounds k = 0..T-1 visit item
        // chunks sj, sj-1, ... (mod T), each from its last item down, where
        // sj is the body's partner chunk: one cyclic descending pass over the
        // items from the last item of chunk sj
        if (lane >= 3 && lane < 6) {
            const int c = lane - 3;
            const int sj = ((il + 1) * T - 1) / n;
            int q = (sj + 1) * nit / T - 1;
            for (int m = 0; m < nit; ++m) {
                if (q != o0 && q != o1) acc += rf[3 * q + c];
                q = q == 0 ? nit - 1 : q - 1;
            }
        }
        // (b) own items: geopotential (idaccel) first, then ring (jdaccel), mass.cpp:298-307
        for (int kind = 0; kind < 2; ++kind) {
            if ((kind == 0 ? o0 : o1) < 0) continue;
            const double *part = kind == 0 ? partA : partB;
            if (lane >= 3 && lane < 9) {
                const int col = lane - 3;                  // 0-2 acceleration, 3-5 torque
                double sum = 0;
                for (int k = 0; k < T; ++k) sum += part[6 * k + col];
                const int f = (kind == 0 ? (col < 3 ? F_IDACC : F_IDTQ - 3) : (col < 3 ? F_JDACC : F_JDTQ - 3)) + col;
                A.st(f, gi, sum);
                acc += sum;
            }
        }
        // (c) daccel and dtorque with RING_CORRECTION (phi and naccel are written)
        if (ring_i < 0) {
            if (lane >= 3 && lane < 6) A.st(F_DACC + lane - 3, gi, acc);
            else if (lane >= 6 && lane < 9) A.st(F_DTQ + lane - 6, gi, acc);
        } else {
            double v9[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) v9[k] = __shfl(acc, k, 64);
            if (lane == 0) {
                const v3 ga{v9[0], v9[1], v9[2]}, nacc{ringS[5], ringS[6], ringS[7]}, mGL{ringS[2], ringS[3], ringS[4]};
                v3 da{v9[3], v9[4], v9[5]}, dq{v9[6], v9[7], v9[8]};
                da = da - ringS[0] * ((ga + da) + nacc);
                const v3 ptorque = dot(dq, mGL) * mGL;
                v3 dtq = dq - ptorque;
                dtq = ringS[1] * dtq;
                dq = dtq + ptorque;
                A.stv(F_DACC, gi, da);
                A.stv(F_DTQ, gi, dq);
            }
        }
        return;
    }
    if (lane >= 3 && lane < 6) A.st(F_DACC + lane - 3, gi, acc);
    else if (lane >= 6 && lane < 9) A.st(F_DTQ + lane - 6, gi, acc);
    if (sm.nitem == 0 && lane == 11) {                     // FINALIZE_RELATIVITY
        A.st(F_PHI, gi, I.phi * kC2);
        A.stv(F_NACC, gi, A.ldv(F_NACCC, gi) * kC2);
    }
}

// deform + spin fixed point + PREPARE_RELATIVITY, one wavefront per body.
// Accumulators: lane 0 phi, 1-3 naccel, 4-12 the tidal matrix; the
// subtracted terms (phi, tidal matrix) are stored negated and the self pair
// contributes +0, as in k_main_wpb.
constexpr int kWpbDefPitch = 65;

// W waves per body: in each round wave w evaluates partner chunk
// (round * W + w); wave 0 then adds the W chunks in partner order.
// TAIL: the launch carries the non-point item fields in extra workgroups
// (without them the kernel needs no scratch).  Work items with il < 0 are
// padding of an XCD-dealt list.
template <int W>
__device__ __forceinline__ void deform_wpb_body(View d, const int2 *__restrict__ work, int prepare, int bid,
                                                double (*term)[13 * kWpbDefPitch]) {
    const Acc A{d.F, d.nbp};
    const int s = work[bid].x, il = work[bid].y;
    if (il < 0) return;
    const SysMeta sm = d.sys[s];
    const int n = sm.n, off = sm.off, gi = off + il;
    const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
    __shared__ double sums[13];
    __shared__ int again_s;
    __shared__ double wsh[3];
    __shared__ double hsh[19];                             // harmonics and GL, gathered by wave 1
    // constant-time-lag bodies repeat the pair loop for each spin pass: with
    // four wavefronts the spin-independent part of every pair is kept here
    constexpr int kGeo = 10, kGeoN = 512;
    __shared__ double geo[W == 4 ? kGeo * kGeoN : 1];
    const ddv3 ri = A.lddv(F_R, gi), vi = A.lddv(F_V, gi);
    const double sR2 = A.ld(F_SR2, gi), k2 = A.ld(F_K2, gi), td = A.ld(F_TD, gi);
    const bool damped = k2 != 0 && td != 0, ndelay = k2 != 0 && td == 0;
    const int nacc_chains = (k2 != 0) ? 13 : 4;
    v3 w = A.ldv(F_W, gi);
    Harmonics H;
    m3 C, GI;
    int iters_left = 4;
    // the spin-independent part of a pair; when every partner has its own
    // lane (one round), it is formed once and kept across spin passes
    const bool one_round = W > 1 && n <= 64 * W;           // (one wavefront: keep registers for occupancy)
    v3 pr{0, 0, 0}, pv{0, 0, 0};
    double prr2 = 0, ptp_dphi = 0, ptp_dg = 0, ptdf = 0;
    auto pair_geo = [&](int jl) {
        const int gj = off + jl;
        pr = diff(A.lddv(F_R, gj), ri);
        prr2 = 1 / smax(dot(pr, pr), sR2 + A.ld(F_SR2, gj));
        const double rr = sqrt(prr2);
        ptp_dphi = rr * A.ld(F_GM, gj);
        ptp_dg = prr2 * ptp_dphi;
        if (damped) {
            pv = diff(A.lddv(F_V, gj), vi);
            ptdf = A.ld(F_TDF, gj);
        }
    };
    if (one_round) {
        const int jl = 64 * wv + lane;
        if (jl < n && jl != il) pair_geo(jl);
    }
    bool first_pass = true;
    for (;;) {
        double acc = (wv == 0 && lane < 4 && !first_pass) ? sums[lane] : 0.0;
        for (int t0 = 0; t0 < n; t0 += 64 * W) {
            const int jl = t0 + 64 * wv + lane;
            double t[13];
#pragma unroll
            for (int q = 0; q < 13; ++q) t[q] = 0.0;
            if (jl < n && jl != il) {
                if (!one_round) {
                    if (W == 4 && damped && n <= kGeoN && !first_pass) {
                        pr = v3{geo[jl], geo[kGeoN + jl], geo[2 * kGeoN + jl]};
                        prr2 = geo[3 * kGeoN + jl]; ptp_dphi = geo[4 * kGeoN + jl]; ptp_dg = geo[5 * kGeoN + jl];
                        pv = v3{geo[6 * kGeoN + jl], geo[7 * kGeoN + jl], geo[8 * kGeoN + jl]};
                        ptdf = geo[9 * kGeoN + jl];
                    } else {
                        pair_geo(jl);
                        if (W == 4 && damped && n <= kGeoN) {
                            geo[jl] = pr.x; geo[kGeoN + jl] = pr.y; geo[2 * kGeoN + jl] = pr.z;
                            geo[3 * kGeoN + jl] = prr2; geo[4 * kGeoN + jl] = ptp_dphi; geo[5 * kGeoN + jl] = ptp_dg;
                            geo[6 * kGeoN + jl] = pv.x; geo[7 * kGeoN + jl] = pv.y; geo[8 * kGeoN + jl] = pv.z;
                            geo[9 * kGeoN + jl] = ptdf;
                        }
                    }
                }
                const v3 r = pr;
                const double rr2 = prr2, tp_dphi = ptp_dphi, tp_dg = ptp_dg;
                const v3 na = tp_dg * r;
                t[0] = -tp_dphi;
                t[1] = na.x; t[2] = na.y; t[3] = na.z;
                if (ndelay) {
                    const double fmj = -3 * tp_dg;
                    const m3 tm = (fmj * k2) * subdiag(outer(rr2 * r, r), 1.0 / 3);
                    const double *tp = &tm.x.x;
#pragma unroll
                    for (int q = 0; q < 9; ++q) t[4 + q] = -tp[q];
                } else if (damped) {
                    const v3 v = pv;
                    const double fmj = -3 * tp_dg;
                    v3 dw = rr2 * cross(r, cross(w, r) - v);
                    double dwt = normalize(dw);
                    dwt *= td * ptdf;
                    const double ecwt2 = 1 + 4 * dwt * dwt,
                                 secwt2 = sqrt(ecwt2),
                                 x0 = sqrt((1 + secwt2) / (2 * ecwt2)),
                                 y0 = dwt * sqrt(2 / (ecwt2 * (1 + secwt2))),
                                 z02 = dwt * dwt * (2 / (ecwt2 + secwt2));
                    const v3 dr = x0 * r + y0 * cross(dw, r);
                    m3 tm = outer(rr2 * dr, dr) - outer(z02 * dw, dw);
                    tm = (fmj * k2) * adddiag(tm, z02 - 1.0 / 3);
                    const double *tp = &tm.x.x;
#pragma unroll
                    for (int q = 0; q < 9; ++q) t[4 + q] = -tp[q];
                }
            }
            if (t0 + 64 * wv < n) {
#pragma unroll
                for (int q = 0; q < 13; ++q) term[wv][q * kWpbDefPitch + lane] = t[q];
            }
            __syncthreads();
            if (wv == 0 && lane < nacc_chains && (first_pass || lane >= 4)) {
                const int nw = min(W, (n - t0 + 63) / 64);
                for (int u = 0; u < nw; ++u) chain_add64(term[u] + lane * kWpbDefPitch, acc);
            }
            // the body's static harmonics in the world frame and GL (neither
            // changes in this kernel before the spin step reads them), while
            // wave 0 adds the first chains
            if (W > 1 && first_pass && t0 == 0 && wv == 1 && lane == 0) {
                Harmonics h;
                h = harmonics_load(A, gi);
                const ddv3 gl = A.lddv(F_GL, gi);
                hsh[0] = h.k2r; hsh[1] = h.scale; hsh[2] = h.R2; hsh[3] = h.A;
                const double *c = &h.Cw.x.x;
#pragma unroll
                for (int q = 0; q < 9; ++q) hsh[4 + q] = c[q];
                hsh[13] = gl.x.hi; hsh[14] = gl.x.lo; hsh[15] = gl.y.hi; hsh[16] = gl.y.lo;
                hsh[17] = gl.z.hi; hsh[18] = gl.z.lo;
            }
            __syncthreads();
        }
        __syncthreads();
        if (wv == 0 && lane < 13) sums[lane] = acc;        // phi, naccel: the first pass's sums, unchanged
        first_pass = false;
        __syncthreads();
        if (threadIdx.x == 0) {
            const m3 Csum{v3{sums[4], sums[5], sums[6]}, v3{sums[7], sums[8], sums[9]}, v3{sums[10], sums[11], sums[12]}};
            ddv3 GL;
            if (W > 1) {
                if (iters_left == 4) {
                    H.k2r = hsh[0]; H.scale = hsh[1]; H.R2 = hsh[2]; H.A = hsh[3];
                    double *c = &H.Cw.x.x;
#pragma unroll
                    for (int q = 0; q < 9; ++q) c[q] = hsh[4 + q];
                }
                GL = ddv3{dd{hsh[13], hsh[14]}, dd{hsh[15], hsh[16]}, dd{hsh[17], hsh[18]}};
            } else {
                if (iters_left == 4) H = harmonics_load(A, gi);
                GL = A.lddv(F_GL, gi);
            }
            bool brk;
            if (damped) {
                H.apply(Csum, w, C, GI);
                brk = !prepare || update_spin(A, gi, GI, w, GL);
                brk = brk || --iters_left == 0;
            } else {
                int iters = 4;
                for (;;) {
                    H.apply(Csum, w, C, GI);
                    if (!prepare || update_spin(A, gi, GI, w, GL) || --iters == 0) break;
                }
                brk = true;
            }
            again_s = brk ? 0 : 1;
            if (brk) deform_write(A, gi, prepare, vi, sums[0], v3{sums[1], sums[2], sums[3]}, C, GI, w);
            else A.stv(F_W, gi, w);
            wsh[0] = w.x; wsh[1] = w.y; wsh[2] = w.z;
        }
        __syncthreads();
        if (!again_s) break;
        w = v3{wsh[0], wsh[1], wsh[2]};
        __syncthreads();
    }
}

template <int W, bool TAIL>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(W == 1 ? 3 : 1))) void k_deform_wpb(View d, const int2 *__restrict__ work, int prepare, NpTail np) {
    if (TAIL && np_tail(d, np)) return;
    __shared__ double term[W][13 * kWpbDefPitch];
    deform_wpb_body<W>(d, work, prepare, (int)blockIdx.x, term);
}

// The large-batch deformation in one launch: workgroups [0, nwork) take the
// constant-time-lag bodies of `work` (one wavefront each, the longest
// workgroups, dispatched first), [nwork, np.ndef) the four-lane blocks of
// the other bodies, which fill the CUs while the last spin passes run, and
// the rest the item fields.  nwork is a multiple of 8, so the block list
// keeps the XCD placement it was dealt with.  The two kinds share one term
// buffer.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_deform_mixed(
    View d, const int2 *__restrict__ work, int nwork, const int4 *__restrict__ blk, int prepare, NpTail np) {
    if (np_tail(d, np)) return;
    constexpr int kTerm = 13 * kWpbDefPitch > 64 * kDefK4Pitch ? 13 * kWpbDefPitch : 64 * kDefK4Pitch;
    __shared__ double term[kTerm];
    const int b = (int)blockIdx.x;
    if (b < nwork) deform_wpb_body<1>(d, work, prepare, b, reinterpret_cast<double (*)[13 * kWpbDefPitch]>(term));
    else deform_k4_body(d, blk, prepare, b - nwork, term);
}

// ---------------------------------------------------------------------------
// main force, single-thread summation order (mass.cpp:329-374): the pair
// terms of partner p acting on body i that the reference adds as reactions
// while visiting p (ROTATIONAL_TIDAL_DEFORMATION and APPLY_NONPOINT_FORCE)
// are added before i's own partner loop when p < i and after it when p > i.
// ---------------------------------------------------------------------------
// field of a model row (View::san) at partner jl
__device__ __forceinline__ v3 san_at(const View &d, int row, int jl) {
    const double *a = d.san + 3 * ((int64_t)row + jl);
    return v3{a[0], a[1], a[2]};
}

// The terms that outer body p adds to partner i's daccel when it visits i
// (ROTATIONAL_TIDAL_DEFORMATION's reaction, then APPLY_NONPOINT_FORCE of
// p's geopotential and ring): -GM_p dg, +GM_p field_gp, +GM_p field_ring,
// in that order; a model p does not have gives +0 (and is not added).
struct SerialReact { v3 t[3]; bool gp, ring; };
__device__ __forceinline__ SerialReact serial_reaction_terms(const View &d, const Acc &A, int gp_, int il, const JMain &T,
                                                             const IMain &I, bool rows) {
    SerialReact o;
    // outer body p visiting partner i: r = r_i - r_p
    const ddv3 rp{dd{T.rxh, T.rxl}, dd{T.ryh, T.ryl}, dd{T.rzh, T.rzl}};
    PairGeo p;
    p.r = diff(I.r, rp);
    p.rr2 = 1 / smax(dot(p.r, p.r), T.sR2 + I.sR2);
    p.rr = sqrt(p.rr2);
    p.rr3 = p.rr * p.rr2;
    const m3 Cp{v3{T.C[0], T.C[1], T.C[2]}, v3{T.C[3], T.C[4], T.C[5]}, v3{T.C[6], T.C[7], T.C[8]}};
    const v3 dg = (p.rr3 * p.rr2 * T.R2) * quad_field(Cp, p);
    o.t[0] = -(T.GM * dg);                                 // da -= GM dg
    const int gpi = d.np_gp[gp_], rgi = d.np_ring[gp_];
    o.gp = gpi >= 0;
    o.ring = rgi >= 0;
    o.t[1] = o.t[2] = v3{0, 0, 0};
    // the field of p's models at i (r = r_i - r_p): from the rows formed
    // beforehand (rows), or evaluated here
    if (o.gp) o.t[1] = T.GM * (rows ? san_at(d, d.srow_gp[gp_], il) : nonpoint_field_call(d, A, gp_, 0, gpi, p.r));
    if (o.ring) o.t[2] = T.GM * (rows ? san_at(d, d.srow_ring[gp_], il) : nonpoint_field_call(d, A, gp_, 1, rgi, p.r));
    return o;
}

__device__ __forceinline__ void serial_reaction(const View &d, const Acc &A, int gp_, int gi, int il, const JMain &T,
                                                const IMain &I, v3 &da, bool rows) {
    const SerialReact o = serial_reaction_terms(d, A, gp_, il, T, I, rows);
    da += o.t[0];
    if (o.gp) da += o.t[1];
    if (o.ring) da += o.t[2];
}

template <int BS>
__global__ __launch_bounds__(BS) void k_main_serial(View d, const int2 *__restrict__ blk) {
    const Acc A{d.F, d.nbp};
    const int s = blk[blockIdx.x].x, i0 = blk[blockIdx.x].y;
    const SysMeta sm = d.sys[s];
    const int n = sm.n, off = sm.off;
    const int il = i0 + (int)threadIdx.x;
    const bool valid = il < n;
    const int gi = off + (valid ? il : 0);
    __shared__ JMain tile[BS];
    const IMain I = load_imain(A, gi);
    const int my_gp = d.np_gp[gi], my_ring = d.np_ring[gi];
    v3 ga{0, 0, 0}, da{0, 0, 0}, dq{0, 0, 0};
    double mind = A.ld(F_MIND, gi), maxi = A.ld(F_MAXI, gi);
    // phase 0: reactions from partners before i; 1: own loop; 2: reactions after i
    for (int phase = 0; phase < 3; ++phase) {
        for (int t0 = 0; t0 < n; t0 += BS) {
            __syncthreads();
            if (t0 + (int)threadIdx.x < n) load_jmain(A, tile[threadIdx.x], off + t0 + (int)threadIdx.x);
            __syncthreads();
            if (!valid) continue;
            const int tn = min(BS, n - t0);
            for (int q = 0; q < tn; ++q) {
                const int jl = t0 + q;
                if (jl == il) continue;
                const JMain &T = tile[q];
                if (phase == 1) {
                    const PairGeo p = relativity(I, T, ga);
                    mind = checked_max_into(mind, p.rr);
                    maxi = checked_max_into(maxi, p.tp_dg);
                    // ROTATIONAL_TIDAL_DEFORMATION, mass.impl:111-116
                    const v3 dg = (p.rr3 * p.rr2 * I.R2) * quad_field(I.C, p);
                    da += T.GM * dg;
                    dq += T.GM * cross(p.r, dg);
                    lt_rad(I, T, p, da, dq);
                    if (my_gp >= 0) {                      // APPLY_NONPOINT_FORCE, mass.impl:127-130
                        const v3 an = d.san ? san_at(d, d.srow_gp[gi], jl) : nonpoint_field_call(d, A, gi, 0, my_gp, p.r);
                        da -= T.GM * an;
                        dq -= T.GM * cross(p.r, an);
                    }
                    if (my_ring >= 0) {
                        const v3 an = d.san ? san_at(d, d.srow_ring[gi], jl) : nonpoint_field_call(d, A, gi, 1, my_ring, p.r);
                        da -= T.GM * an;
                        dq -= T.GM * cross(p.r, an);
                    }
                } else if ((phase == 0) == (jl < il)) {
                    serial_reaction(d, A, off + jl, gi, il, T, I, da, d.san != nullptr);
                }
            }
        }
    }
    if (!valid) return;
    A.stv(F_GACC, gi, ga);
    A.st(F_MIND, gi, mind);
    A.st(F_MAXI, gi, maxi);
    // FINALIZE_RELATIVITY and RING_CORRECTION (mass.impl:133-152)
    finalize_body(d, A, gi, da, dq);
}

// ---------------------------------------------------------------------------
// non-point items, thread-pool order (mass.cpp:158-214, 253-307)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_np_field(View d, int nrows) {
    // one thread per (item, partner): an = field of the item at r_j - r_p
    const Acc A{d.F, d.nbp};
    const int row = blockIdx.y;
    if (row >= nrows) return;
    np_field_one(d, A, row, blockIdx.x * blockDim.x + threadIdx.x);
}

// one wavefront per (item, round k): the item's partial sums over the slice
// it visits in round k (tpmi of accel_nonpoint).  Lanes form the partners'
// terms (subtracted terms stored negated, so a - b becomes a + (-b), the same
// value); lane c < 6 adds chain c in partner order (td x, y, z, tq x, y, z).
// The self pair and lanes past the slice add +0, an exact identity for sums
// that start at +0.  bid = item row * maxT + round; term holds 64 x 7 doubles.
__device__ __forceinline__ void np_item_body(View d, int nrows, int maxT, int bid, double *term) {
    const Acc A{d.F, d.nbp};
    const int k = bid % maxT, it_i = bid / maxT;
    if (it_i >= nrows) return;
    const NPItem it = d.items[it_i];
    const SysMeta sm = d.sys[it.sys];
    const int T = sm.T;
    if (k >= T) return;
    int kk = it.chain + k;
    if (kk >= T) kk -= T;
    const int j0 = kk * sm.n / T, j1 = (kk + 1) * sm.n / T;
    const int lane = (int)threadIdx.x;
    const ddv3 rp = A.lddv(F_R, it.body);
    double acc = 0;
    for (int base = j0; base < j1; base += 64) {
        const int jl = base + lane;
        double t[6] = {0, 0, 0, 0, 0, 0};
        if (jl < j1 && sm.off + jl != it.body) {
            const int gj = sm.off + jl;
            const v3 r = diff(A.lddv(F_R, gj), rp);
            const double *a = d.an + 3 * ((int64_t)it.an_off + jl);
            const v3 an{a[0], a[1], a[2]};
            const double GMj = A.ld(F_GM, gj);
            const v3 ta = GMj * an, tq = GMj * cross(r, an);
            t[0] = -ta.x; t[1] = -ta.y; t[2] = -ta.z;
            t[3] = -tq.x; t[4] = -tq.y; t[5] = -tq.z;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) term[lane * 7 + c] = t[c];
        __syncthreads();
        if (lane < 6) {
            const int cnt = min(64, j1 - base);
            for (int q = 0; q < cnt; ++q) acc += term[q * 7 + lane];
        }
        __syncthreads();
    }
    if (lane < 6) d.part[6 * ((int64_t)it_i * maxT + k) + lane] = acc;
}

__global__ __launch_bounds__(64) void k_np_item(View d, int nrows, int maxT) {
    __shared__ double term[64 * 7];
    np_item_body(d, nrows, maxT, (int)blockIdx.x, term);
}

// The large-batch main force with the items' partial sums in extra
// workgroups: [0, nblk) the four-lane blocks (in the XCD placement they were
// dealt with), then one wavefront per (item row, round).  The item sums read
// only positions, masses and the item fields formed in the deformation
// launch, so they fill the CUs while the last blocks run.  One term buffer
// serves both.
static_assert(kK4Pitch >= 7, "the item sums use 7 doubles per lane of the term buffer");
__global__ __launch_bounds__(64) void k_main_k4_items(View d, const int2 *__restrict__ blk, int nblk, int nrows, int maxT) {
    __shared__ double term[64 * kK4Pitch];
    const int b = (int)blockIdx.x;
    if (b < nblk) main_k4_body(d, blk, b, term);
    else np_item_body(d, nrows, maxT, b - nblk, term);
}

constexpr int kRecvStaged = 64;          // items per system staged in LDS
__global__ __launch_bounds__(kBS) void k_np_recv(View d, const int2 *__restrict__ blk, int maxT) {
    // one thread per body of a system with items: receiver terms in round
    // order, then the body's own item sums, finalisation and ring correction.
    // The system's items (body, row offset, GM of the body) are staged in LDS
    // once per workgroup.
    const Acc A{d.F, d.nbp};
    const int s = blk[blockIdx.x].x, i0 = blk[blockIdx.x].y;
    const SysMeta sm = d.sys[s];
    const int T = sm.T, nit = sm.nitem, n = sm.n;
    __shared__ int s_body[kRecvStaged], s_an[kRecvStaged], s_kind[kRecvStaged];
    __shared__ double s_gm[kRecvStaged];
    const bool staged = nit <= kRecvStaged;
    if (staged && (int)threadIdx.x < nit) {
        const NPItem it = d.items[sm.item0 + threadIdx.x];
        s_body[threadIdx.x] = it.body;
        s_an[threadIdx.x] = it.an_off;
        s_kind[threadIdx.x] = it.kind;
        s_gm[threadIdx.x] = A.ld(F_GM, it.body);
    }
    __syncthreads();
    const int jl = i0 + (int)threadIdx.x;
    if (jl >= n) return;
    const int gj = sm.off + jl;
    v3 da = A.ldv(F_DACC, gj), dq = A.ldv(F_DTQ, gj);
    // rounds k = 0..T-1 visit item chunks sj, sj-1, ... (mod T), each from
    // its last item down (sj: the body's partner chunk): one cyclic
    // descending pass over the items from the last item of chunk sj.  The
    // item sequence is known in advance, so the fields are loaded ahead of
    // the ordered additions.
    const int sj = ((jl + 1) * T - 1) / n;
    int q = (sj + 1) * nit / T - 1;
#pragma unroll 4
    for (int m = 0; m < nit; ++m) {
        int body, an_off;
        double gm;
        if (staged) {
            body = s_body[q]; an_off = s_an[q]; gm = s_gm[q];
        } else {
            const NPItem it = d.items[sm.item0 + q];
            body = it.body; an_off = it.an_off; gm = A.ld(F_GM, it.body);
        }
        const double *a = d.an + 3 * ((int64_t)an_off + jl);
        const v3 an{a[0], a[1], a[2]};
        if (body != gj) da += gm * an;
        q = q == 0 ? nit - 1 : q - 1;
    }
    // own items: idaccel (geopotential) then jdaccel (ring), mass.cpp:298-307
    for (int kind = 0; kind < 2; ++kind) {
        for (int q = 0; q < nit; ++q) {
            int body, knd;
            if (staged) {
                body = s_body[q]; knd = s_kind[q];
            } else {
                const NPItem it = d.items[sm.item0 + q];
                body = it.body; knd = it.kind;
            }
            if (body != gj || knd != kind) continue;
            v3 sa{0, 0, 0}, sq{0, 0, 0};
            for (int k = 0; k < T; ++k) {
                const double *p = d.part + 6 * ((int64_t)(sm.item0 + q) * maxT + k);
                sa += v3{p[0], p[1], p[2]};
                sq += v3{p[3], p[4], p[5]};
            }
            A.stv(kind == 0 ? F_IDACC : F_JDACC, gj, sa);
            A.stv(kind == 0 ? F_IDTQ : F_JDTQ, gj, sq);
            da += sa;
            dq += sq;
        }
    }
    finalize_body(d, A, gj, da, dq);
}

// ---------------------------------------------------------------------------
// RK12 stage: store f[k-1], combine, update state (RungeKutta.impl:113-186)
// ---------------------------------------------------------------------------
// The stage update splits into two independent halves: the translational
// one (r, v, the stage's acceleration, GM(t), exJ2(t)) and the rotational one
// (GL, the frame s, w, torque, Egrad).  Neither reads a field the other
// writes, so k_stage runs them in different workgroups.
__device__ __forceinline__ void stage_translational(const View &d, const Acc &A, int g, int k, double dt, double tadd) {
    const ddv3 v = A.lddv(F_V, g);
    if (k == 1) {
        A.stddv(F_X0R, g, A.lddv(F_R, g));
        A.stddv(F_X0V, g, v);
    }
    const v3 a_now = (A.ldv(F_GACC, g) + A.ldv(F_DACC, g)) + A.ldv(F_NACC, g);
    {
        const int fb = F_STAGE + (k - 1) * kStageFields;
        A.stddv(fb, g, v);
        A.stv(fb + 6, g, a_now);
    }
    const ddv3 r0 = A.lddv(F_X0R, g), v0 = A.lddv(F_X0V, g);
    const double gm0 = A.ld(F_GM0, g), dgm = A.ld(F_DGM, g), dj2 = A.ld(F_DJ2, g);
    const double t = d.t0[d.bsys[g]] + tadd;
    ddv3 cv = mkddv(v3{0, 0, 0});
    v3 ca{0, 0, 0};
    const int base = k * (k - 1) / 2;
    const int e1 = c_row0[k + 1];
#pragma unroll 4
    for (int e = c_row0[k]; e < e1; ++e) {              // non-zero columns, ascending
        const int j = c_nzj[e];
        const double2 c = c_rk[base + j];
        const dd ckj{c.x, c.y};
        const double fc = c.x + c.y;
        const int fb = F_STAGE + j * kStageFields;
        const bool now = j == k - 1;                    // this stage's record, still in registers
        cv = add(cv, scale(ckj, now ? v : A.lddv(fb, g)));
        ca += fc * (now ? a_now : A.ldv(fb + 6, g));
    }
    A.stddv(F_R, g, add(r0, scale(mkdd(dt), cv)));
    A.stddv(F_V, g, add(v0, mkddv(dt * ca)));
    A.st(F_GM, g, gm0 + dgm * t);
    A.st(F_EXJ2, g, dj2 * t);
}

__device__ __forceinline__ void stage_rotational(const Acc &A, int g, int k, double dt, double dt_k) {
    if (k == 1) {
        A.stddv(F_X0GL, g, A.lddv(F_GL, g));
        A.stddm(F_X0S, g, A.lddm(F_S, g));
        A.stv(F_X0W, g, A.ldv(F_W, g));
        A.stv(F_EGRAD, g, A.ldv(F_DTQ, g));
    }
    const v3 w_now = A.ldv(F_W, g), tq_now = A.ldv(F_DTQ, g);
    {
        const int fb = F_STAGE + (k - 1) * kStageFields;
        A.stv(fb + 9, g, w_now);
        A.stv(fb + 12, g, tq_now);
    }
    v3 egrad = A.ldv(F_EGRAD, g);
    if (k == 25) {
        egrad = egrad_k25(A.ldm(F_GI, g), todbl(A.lddv(F_GL, g)), w_now, tq_now, egrad, dt);
        A.stv(F_EGRAD, g, egrad);
    }
    const ddv3 gl0 = A.lddv(F_X0GL, g);
    const v3 w0 = A.ldv(F_X0W, g);
    ddm3 sm = A.lddm(F_X0S, g);
    v3 cw{0, 0, 0}, ct{0, 0, 0};
    const int base = k * (k - 1) / 2;
    const int e1 = c_row0[k + 1];
#pragma unroll 4
    for (int e = c_row0[k]; e < e1; ++e) {              // non-zero columns, ascending
        const int j = c_nzj[e];
        const double2 c = c_rk[base + j];
        const double fc = c.x + c.y;
        const int fb = F_STAGE + j * kStageFields;
        const bool now = j == k - 1;                    // this stage's record, still in registers
        cw += fc * (now ? w_now : A.ldv(fb + 9, g));
        ct += fc * (now ? tq_now : A.ldv(fb + 12, g));
    }
    const v3 w = cw / dt_k;
    const v3 dw = (2.0 / 3) * (w - w0);
    sm = add(sm, rotation_matrix(w - dw, dt * dt_k / 2, DevSin{}) * todbl(sm));
    sm = add(sm, rotation_matrix(w + dw, dt * dt_k / 2, DevSin{}) * todbl(sm));
    A.stddv(F_GL, g, add(gl0, mkddv(dt * ct)));
    if (k == 25) {
        sm = add(sm, rotation_matrix(egrad, 1.0 / 72, DevSin{}) * todbl(sm));
        sm = orthogonalized(sm);
    }
    A.stddm(F_S, g, sm);
    A.stv(F_W, g, w);
}

// the persistent small-system kernel keeps the two halves of the stage
// update out of line and runs them on different threads
__device__ __noinline__ void stage_translational_ool(const View &d, const Acc &A, int g, int k, double dt, double tadd) {
    stage_translational(d, A, g, k, dt, tadd);
}
__device__ __noinline__ void stage_rotational_ool(const Acc &A, int g, int k, double dt, double dt_k) {
    stage_rotational(A, g, k, dt, dt_k);
}

// Component c (0, 1, 2 = x, y, z) of the translational stage update: the
// same operations as stage_translational on that component of every
// vector (they never mix components), so each thread's results are the
// bits of the one-thread update.
__device__ __forceinline__ void stage_translational_c(const View &d, const Acc &A, int g, int c, int k, double dt,
                                                      double tadd) {
    const dd v = A.ldd(F_V + 2 * c, g);
    if (k == 1) {
        A.std_(F_X0R + 2 * c, g, A.ldd(F_R + 2 * c, g));
        A.std_(F_X0V + 2 * c, g, v);
    }
    const double a_now = (A.ld(F_GACC + c, g) + A.ld(F_DACC + c, g)) + A.ld(F_NACC + c, g);
    {
        const int fb = F_STAGE + (k - 1) * kStageFields;
        A.std_(fb + 2 * c, g, v);
        A.st(fb + 6 + c, g, a_now);
    }
    const dd r0 = A.ldd(F_X0R + 2 * c, g), v0 = A.ldd(F_X0V + 2 * c, g);
    dd cv = mkdd(0.0);
    double ca = 0;
    const int base = k * (k - 1) / 2;
    const int e1 = c_row0[k + 1];
#pragma unroll 4
    for (int e = c_row0[k]; e < e1; ++e) {              // non-zero columns, ascending
        const int j = c_nzj[e];
        const double2 ck = c_rk[base + j];
        const dd ckj{ck.x, ck.y};
        const double fc = ck.x + ck.y;
        const int fb = F_STAGE + j * kStageFields;
        const bool now = j == k - 1;                    // this stage's record, still in registers
        cv = add(cv, mul(ckj, now ? v : A.ldd(fb + 2 * c, g)));
        ca += fc * (now ? a_now : A.ld(fb + 6 + c, g));
    }
    A.std_(F_R + 2 * c, g, add(r0, mul(mkdd(dt), cv)));
    A.std_(F_V + 2 * c, g, add(v0, mkdd(dt * ca)));
    if (c == 0) {
        const double t = d.t0[d.bsys[g]] + tadd;
        A.st(F_GM, g, A.ld(F_GM0, g) + A.ld(F_DGM, g) * t);
        A.st(F_EXJ2, g, A.ld(F_DJ2, g) * t);
    }
}

// istep != nullptr: the step index comes from device memory (replayed graphs)
// and the stage time offset dt * (dt_k + i_step) is formed here, with the
// same two operations as on the host.
// Workgroups [0, nblk) take the translational half, [nblk, 2 nblk) the
// rotational half of the same 64 bodies; wavefront c of a workgroup takes
// vector component c, so the long dependent sums over the stage records run
// on three lanes per body instead of one.  In the rotational half the
// component sums of the spin meet in LDS and wavefront 0 then forms the
// rotation of the frame (stage_rotational's operations, in its order).
__global__ __launch_bounds__(192) void k_stage(View d, int k, double dt, double dt_k, double tadd,
                                               const int64_t *__restrict__ istep) {
    const Acc A{d.F, d.nbp};
    const int nblk = (d.nb + 63) / 64;
    const bool rot = (int)blockIdx.x >= nblk;
    const int lane = (int)threadIdx.x & 63, c = (int)threadIdx.x >> 6;
    const int g = ((int)blockIdx.x - (rot ? nblk : 0)) * 64 + lane;
    const bool valid = g < d.nb;
    if (istep) tadd = dt * (dt_k + (double)*istep);
    if (!rot) {
        if (valid) stage_translational_c(d, A, g, c, k, dt, tadd);
        return;
    }
    __shared__ double cw_s[3][64];
    const int gg = valid ? g : 0;
    dd gl_new{0, 0};
    double cw = 0;
    // wavefront 0 at the last stage: the inputs of the energy-gradient
    // correction, read before any wavefront stores this stage's GL
    m3 GI = mzero();
    v3 gl_old{0, 0, 0}, w_full{0, 0, 0}, tq_full{0, 0, 0}, egrad{0, 0, 0};
    if (valid) {
        if (k == 1) {
            A.std_(F_X0GL + 2 * c, gg, A.ldd(F_GL + 2 * c, gg));
            A.stddv(F_X0S + 6 * c, gg, A.lddv(F_S + 6 * c, gg));
            A.st(F_X0W + c, gg, A.ld(F_W + c, gg));
            A.st(F_EGRAD + c, gg, A.ld(F_DTQ + c, gg));
        }
        const double w_now = A.ld(F_W + c, gg), tq_now = A.ld(F_DTQ + c, gg);
        {
            const int fb = F_STAGE + (k - 1) * kStageFields;
            A.st(fb + 9 + c, gg, w_now);
            A.st(fb + 12 + c, gg, tq_now);
        }
        if (c == 0 && k == 25) {
            GI = A.ldm(F_GI, gg);
            gl_old = todbl(A.lddv(F_GL, gg));
            w_full = A.ldv(F_W, gg);
            tq_full = A.ldv(F_DTQ, gg);
            egrad = A.ldv(F_EGRAD, gg);
        }
        const dd gl0 = A.ldd(F_X0GL + 2 * c, gg);
        double ct = 0;
        const int base = k * (k - 1) / 2;
        const int e1 = c_row0[k + 1];
#pragma unroll 4
        for (int e = c_row0[k]; e < e1; ++e) {          // non-zero columns, ascending
            const int j = c_nzj[e];
            const double2 ck = c_rk[base + j];
            const double fc = ck.x + ck.y;
            const int fb = F_STAGE + j * kStageFields;
            const bool now = j == k - 1;
            cw += fc * (now ? w_now : A.ld(fb + 9 + c, gg));
            ct += fc * (now ? tq_now : A.ld(fb + 12 + c, gg));
        }
        gl_new = add(gl0, mkdd(dt * ct));
    }
    cw_s[c][lane] = cw;
    __syncthreads();
    if (!valid) return;
    A.std_(F_GL + 2 * c, gg, gl_new);
    if (c != 0) return;
    if (k == 25) {
        egrad = egrad_k25(GI, gl_old, w_full, tq_full, egrad, dt);
        A.stv(F_EGRAD, gg, egrad);
    }
    // at the first stage the saved start values equal the current ones
    const v3 w0 = A.ldv(k == 1 ? F_W : F_X0W, gg);
    ddm3 sm = A.lddm(k == 1 ? F_S : F_X0S, gg);
    const v3 w = v3{cw_s[0][lane], cw_s[1][lane], cw_s[2][lane]} / dt_k;
    const v3 dw = (2.0 / 3) * (w - w0);
    sm = add(sm, rotation_matrix(w - dw, dt * dt_k / 2, DevSin{}) * todbl(sm));
    sm = add(sm, rotation_matrix(w + dw, dt * dt_k / 2, DevSin{}) * todbl(sm));
    if (k == 25) {
        sm = add(sm, rotation_matrix(egrad, 1.0 / 72, DevSin{}) * todbl(sm));
        sm = orthogonalized(sm);
    }
    A.stddm(F_S, gg, sm);
    A.stv(F_W, gg, w);
}

__global__ void k_step_count(int64_t *istep) {
    if (threadIdx.x == 0) *istep += 1;
}

// ---------------------------------------------------------------------------
// whole-system-in-one-workgroup path: every body of the system sits in one
// LDS tile, so deform and main force need no further tile traffic
// ---------------------------------------------------------------------------
// Deformation of a small system by the whole workgroup, pass by pass: in
// each pass the threads form the terms of the pairs (i, q) of the bodies
// still iterating -- -phi, naccel and the subtracted tidal matrix of
// deform_by (mass.cpp:78-103), the last one at the body's current spin for a
// constant-time-lag body -- into LDS, thirteen threads per body add them in
// partner order, and the body's own thread runs UPDATE_HARMONICS and
// UPDATE_ANGULAR_VELOCITY.  A body without delayed tides sums once and
// iterates its spin on those sums (its terms do not depend on the spin);
// a constant-time-lag body re-sums its tidal matrix at every pass, as the
// reference does, up to MAX_ANGULAR_VELOCITY_ITER passes (phi and naccel do
// not change between passes).
struct SmallDefShared {
    double w[kSmallMax * 3], k2[kSmallMax], td[kSmallMax];
    int act[kSmallMax], list[kSmallMax], na;
};

__device__ __forceinline__ void small_deform_terms(const JDeform &Ti, const JDeform &T, double k2, double td, v3 w,
                                                   bool first, double *o) {
    const ddv3 ri{dd{Ti.rxh, Ti.rxl}, dd{Ti.ryh, Ti.ryl}, dd{Ti.rzh, Ti.rzl}};
    const ddv3 rj{dd{T.rxh, T.rxl}, dd{T.ryh, T.ryl}, dd{T.rzh, T.rzl}};
    const v3 r = diff(rj, ri);
    const double rr2 = 1 / smax(dot(r, r), Ti.sR2 + T.sR2);
    const double rr = sqrt(rr2);
    const double tp_dphi = rr * T.GM;
    const double tp_dg = rr2 * tp_dphi;
    if (first) {
        const v3 na = tp_dg * r;
        o[0] = -tp_dphi;                                   // phi -= tp_dphi
        o[1] = na.x; o[2] = na.y; o[3] = na.z;             // naccel += tp_dg r
    }
    m3 X = mzero();                                        // Csum -= X
    if (k2 != 0 && td != 0) {                              // DAMPED_TIDAL_DEFORMATION_MATRIX
        const ddv3 vi{dd{Ti.vxh, Ti.vxl}, dd{Ti.vyh, Ti.vyl}, dd{Ti.vzh, Ti.vzl}};
        const ddv3 vj{dd{T.vxh, T.vxl}, dd{T.vyh, T.vyl}, dd{T.vzh, T.vzl}};
        const v3 v = diff(vj, vi);
        const double fmj = -3 * tp_dg;
        v3 dw = rr2 * cross(r, cross(w, r) - v);
        double dwt = normalize(dw);
        dwt *= td * T.tdf;
        const double ecwt2 = 1 + 4 * dwt * dwt,
                     secwt2 = sqrt(ecwt2),
                     x0 = sqrt((1 + secwt2) / (2 * ecwt2)),
                     y0 = dwt * sqrt(2 / (ecwt2 * (1 + secwt2))),
                     z02 = dwt * dwt * (2 / (ecwt2 + secwt2));
        const v3 dr = x0 * r + y0 * cross(dw, r);
        m3 t = outer(rr2 * dr, dr) - outer(z02 * dw, dw);
        t = adddiag(t, z02 - 1.0 / 3);
        X = (fmj * k2) * t;
    } else if (k2 != 0) {                                  // _NDELAY
        const double fmj = -3 * tp_dg;
        X = (fmj * k2) * subdiag(outer(rr2 * r, r), 1.0 / 3);
    }
    const double *x = &X.x.x;
#pragma unroll
    for (int c = 0; c < 9; ++c) o[4 + c] = -x[c];
}

// called by every thread of the workgroup of NT threads (it synchronises)
template <int NT>
__device__ void deform_small(const Acc &A, int off, int n, const JDeform *tile, double *terms, double *sums,
                             SmallDefShared &sh) {
    const int tid = (int)threadIdx.x;
    const bool body = tid < n;
    const int gi = off + (body ? tid : 0);
    // the body's own state, held by its thread across the passes
    ddv3 vi{}, GL{};
    Harmonics H{};
    v3 w{0, 0, 0}, nacc{0, 0, 0};
    double phi = 0, k2 = 0, td = 0;
    m3 C = mzero(), GI = mzero();
    if (body) {
        vi = A.lddv(F_V, gi);
        k2 = A.ld(F_K2, gi);
        td = A.ld(F_TD, gi);
        GL = A.lddv(F_GL, gi);
        H = harmonics_load(A, gi);
        w = A.ldv(F_W, gi);
        sh.w[3 * tid] = w.x; sh.w[3 * tid + 1] = w.y; sh.w[3 * tid + 2] = w.z;
        sh.k2[tid] = k2;
        sh.td[tid] = td;
        sh.act[tid] = 1;
        sh.list[tid] = tid;
    }
    if (tid == 0) sh.na = n;
    __syncthreads();
    const int m = n - 1;                                   // pairs per body
    const int bpr = m > 0 ? min(NT / m, NT / 13) : 1;      // bodies per round
    for (int pass = 0; pass < 4; ++pass) {
        const int na = sh.na;
        if (na == 0) break;
        const bool first = pass == 0;
        for (int b0 = 0; b0 < na && m > 0; b0 += bpr) {
            const int nb = min(bpr, na - b0);
            if (tid < nb * m) {
                const int il = sh.list[b0 + tid / m], qq = tid - (tid / m) * m;
                const int q = qq < il ? qq : qq + 1;
                small_deform_terms(tile[il], tile[q], sh.k2[il], sh.td[il],
                                   v3{sh.w[3 * il], sh.w[3 * il + 1], sh.w[3 * il + 2]}, first, terms + tid * 13);
            }
            __syncthreads();
            if (tid < 13 * nb) {
                const int j = tid / 13, c = tid - j * 13;
                const int il = sh.list[b0 + j];
                if (first || c >= 4) {
                    const double *P = terms + j * m * 13 + c;
                    double acc = 0;
                    for (int qq = 0; qq < m; ++qq) acc += P[qq * 13];
                    sums[il * 13 + c] = acc;
                }
            }
            __syncthreads();
        }
        if (m == 0 && body) {                              // no partner: the sums stay +0
            for (int c = 0; c < 13; ++c) sums[tid * 13 + c] = 0.0;
        }
        __syncthreads();
        if (body && sh.act[tid]) {
            const double *S = sums + tid * 13;
            if (first) {
                phi = S[0];
                nacc = v3{S[1], S[2], S[3]};
            }
            const m3 Csum{v3{S[4], S[5], S[6]}, v3{S[7], S[8], S[9]}, v3{S[10], S[11], S[12]}};
            if (!(k2 != 0 && td != 0)) {                   // the spin iterates on the same sums
                int iters = 4;                             // MAX_ANGULAR_VELOCITY_ITER
                for (;;) {
                    H.apply(Csum, w, C, GI);
                    if (update_spin(A, gi, GI, w, GL) || --iters == 0) break;
                }
                sh.act[tid] = 0;
            } else {
                H.apply(Csum, w, C, GI);
                const bool brk = update_spin(A, gi, GI, w, GL);
                sh.act[tid] = (brk || pass == 3) ? 0 : 1;
                sh.w[3 * tid] = w.x; sh.w[3 * tid + 1] = w.y; sh.w[3 * tid + 2] = w.z;
            }
        }
        __syncthreads();
        if (tid == 0) {
            int k = 0;
            for (int i = 0; i < n; ++i)
                if (sh.act[i]) sh.list[k++] = i;
            sh.na = k;
        }
        __syncthreads();
    }
    if (body) deform_write(A, gi, 1, vi, phi, nacc, C, GI, w);
}

// Main force of a small system in the single-thread order, by the whole
// workgroup: each thread forms every term of one ordered pair (i, q) -- the
// own-loop terms of i's partner q and the reaction terms of outer body q on
// i -- into LDS; then ten threads per body add them in the order of the
// sequential loop (mass.cpp:329-374): daccel takes the reactions of q < i,
// the own-loop terms of every q, the reactions of q > i; dtorque and gaccel
// the own-loop terms; the running maxima go over the partners.  A body's
// pairs fall in one round, so its sums are never split.  Terms of a model
// the body does not have are +0, an exact identity in these sums.
constexpr int kSmallTerm = 44;                         // doubles per pair
// pair slots: 0-2 G; 3-20 daccel own terms (GM dg, -LT own spin, +LT
// partner spin, radiation, -GM field_gp, -GM field_ring); 21-32 dtorque
// own terms (GM r x dg, -LT torque, -GM r x field_gp, -GM r x field_ring);
// 33-41 daccel reaction terms (-GM_q dg_q, +GM_q field_gp, +GM_q field_ring);
// 42 1/r, 43 GM/r^3
__device__ void small_pair_terms(const View &d, const Acc &A, int off, int il, int q, const JMain &T, double *o) {
    const int gi = off + il;
    const IMain I = load_imain(A, gi);
    const int my_gp = d.np_gp[gi], my_ring = d.np_ring[gi];
    v3 unused{0, 0, 0};
    const PairGeo p = relativity(I, T, unused);
    v3 ta[6], tq[4];
    const v3 dg = (p.rr3 * p.rr2 * I.R2) * quad_field(I.C, p);
    ta[0] = T.GM * dg;
    tq[0] = T.GM * cross(p.r, dg);
    // LENSE_THIRRING and RADIATION_PRESSURE as lt_rad forms them
    const double rcr32 = 2 * T.GM / kC * p.rr3;
    const v3 GLb = cross(I.GL, p.b);
    ta[1] = -(rcr32 * (GLb - (3 * dot(GLb, p.r) * p.rr2) * p.r));
    tq[1] = -(rcr32 * cross(I.GL, cross(p.r, p.b)));
    const v3 GLj{T.GLx, T.GLy, T.GLz};
    ta[2] = cross(rcr32 * (GLj - (3 * dot(GLj, p.r) * p.rr2) * p.r), p.b);
    ta[3] = (T.lum * I.rR2G4c * p.rr3 / I.GM) * p.r;
    ta[4] = ta[5] = tq[2] = tq[3] = v3{0, 0, 0};
    if (my_gp >= 0) {
        const v3 an = nonpoint_field_call(d, A, gi, 0, my_gp, p.r);
        ta[4] = -(T.GM * an);
        tq[2] = -(T.GM * cross(p.r, an));
    }
    if (my_ring >= 0) {
        const v3 an = nonpoint_field_call(d, A, gi, 1, my_ring, p.r);
        ta[5] = -(T.GM * an);
        tq[3] = -(T.GM * cross(p.r, an));
    }
    const SerialReact r = serial_reaction_terms(d, A, off + q, il, T, I, false);
    o[0] = p.G.x; o[1] = p.G.y; o[2] = p.G.z;
#pragma unroll
    for (int k = 0; k < 6; ++k) { o[3 + 3 * k] = ta[k].x; o[4 + 3 * k] = ta[k].y; o[5 + 3 * k] = ta[k].z; }
#pragma unroll
    for (int k = 0; k < 4; ++k) { o[21 + 3 * k] = tq[k].x; o[22 + 3 * k] = tq[k].y; o[23 + 3 * k] = tq[k].z; }
#pragma unroll
    for (int k = 0; k < 3; ++k) { o[33 + 3 * k] = r.t[k].x; o[34 + 3 * k] = r.t[k].y; o[35 + 3 * k] = r.t[k].z; }
    o[42] = p.rr;
    o[43] = p.tp_dg;
}

// called by every thread of the workgroup of NT threads (it synchronises)
template <int NT>
__device__ void main_serial_small(const View &d, const Acc &A, int off, int n, const JMain *tile, double *terms,
                                  double *sums) {
    const int tid = (int)threadIdx.x;
    if (n > 1) {
        const int m = n - 1;                                // pairs per body
        const int bpr = min(12, NT / m);                    // bodies per round (10 sum threads each)
        for (int b0 = 0; b0 < n; b0 += bpr) {
            const int nb = min(bpr, n - b0);
            if (tid < nb * m) {
                const int il = b0 + tid / m, qq = tid - (tid / m) * m;
                const int q = qq < il ? qq : qq + 1;
                small_pair_terms(d, A, off, il, q, tile[q], terms + tid * kSmallTerm);
            }
            __syncthreads();
            if (tid < 10 * nb) {
                const int il = b0 + tid / 10, c = tid - (tid / 10) * 10;
                const double *P = terms + (il - b0) * m * kSmallTerm;   // this body's pairs, q ascending
                double acc = 0;
                if (c < 3) {                                // daccel
                    for (int qq = 0; qq < il; ++qq)
                        for (int k = 0; k < 3; ++k) acc += P[qq * kSmallTerm + 33 + 3 * k + c];
                    for (int qq = 0; qq < m; ++qq)
                        for (int k = 0; k < 6; ++k) acc += P[qq * kSmallTerm + 3 + 3 * k + c];
                    for (int qq = il; qq < m; ++qq)
                        for (int k = 0; k < 3; ++k) acc += P[qq * kSmallTerm + 33 + 3 * k + c];
                } else if (c < 6) {                         // dtorque
                    for (int qq = 0; qq < m; ++qq)
                        for (int k = 0; k < 4; ++k) acc += P[qq * kSmallTerm + 21 + 3 * k + (c - 3)];
                } else if (c < 9) {                         // gaccel
                    for (int qq = 0; qq < m; ++qq) acc += P[qq * kSmallTerm + (c - 6)];
                } else {                                    // running maxima
                    const int gi = off + il;
                    double mind = A.ld(F_MIND, gi), maxi = A.ld(F_MAXI, gi);
                    for (int qq = 0; qq < m; ++qq) {
                        mind = checked_max_into(mind, P[qq * kSmallTerm + 42]);
                        maxi = checked_max_into(maxi, P[qq * kSmallTerm + 43]);
                    }
                    sums[il * 11 + 10] = maxi;
                    acc = mind;
                }
                sums[il * 11 + c] = acc;
            }
            __syncthreads();
        }
    }
    if (tid < n) {
        const int gi = off + tid;
        const double *S = sums + tid * 11;
        const bool any = n > 1;
        A.stv(F_GACC, gi, any ? v3{S[6], S[7], S[8]} : v3{0, 0, 0});
        if (any) {
            A.st(F_MIND, gi, S[9]);
            A.st(F_MAXI, gi, S[10]);
        }
        finalize_body(d, A, gi, any ? v3{S[0], S[1], S[2]} : v3{0, 0, 0}, any ? v3{S[3], S[4], S[5]} : v3{0, 0, 0});
    }
}

// One workgroup integrates one system (N <= kSmallMax bodies, single-thread
// summation order) for n_steps RK12 steps with no launch per stage: thread
// t < N takes body t, and threads kSmallMax + t the rotational half of the
// stage update of body t (the two halves touch different fields).  With
// `rec`, the partner data (JMain) of every body after each stage's deform is
// recorded: entry e = 25*step + (k-1) of the launch, workgroup b, body t at
// rec[(e*nrec_sys + b)*kSmallMax + t].
template <int NT>
__global__ __launch_bounds__(NT) void k_small(View d, const int *__restrict__ sys_ids, int n_steps,
                                                     double dt, DtK dtk, int64_t i_step0, JMain *rec, int nrec_sys) {
    const Acc A{d.F, d.nbp};
    const int s = sys_ids[blockIdx.x];
    const SysMeta sm = d.sys[s];
    const int n = sm.n, off = sm.off, t = (int)threadIdx.x;
    const bool valid = t < n;
    __shared__ JDeform tdf[kSmallMax];
    __shared__ JMain tmn[kSmallMax];
    __shared__ double terms[NT * kSmallTerm];
    __shared__ double sums[kSmallMax * 13];
    __shared__ SmallDefShared dsh;
    for (int step = 0; step < n_steps; ++step) {
        for (int k = 1; k <= 25; ++k) {
            const double tadd = dt * (dtk.v[k] + (double)(i_step0 + step));
            if (valid) stage_translational_ool(d, A, off + t, k, dt, tadd);
            else if (t >= kSmallMax && t - kSmallMax < n) stage_rotational_ool(A, off + t - kSmallMax, k, dt, dtk.v[k]);
            __syncthreads();
            if (valid) load_jdeform(A, tdf[t], off + t);
            __syncthreads();
            deform_small<NT>(A, off, n, tdf, terms, sums, dsh);
            __syncthreads();
            if (valid) load_jmain(A, tmn[t], off + t);
            __syncthreads();
            main_serial_small<NT>(d, A, off, n, tmn, terms, sums);
            if (rec && valid) {
                const int64_t e = 25 * (int64_t)step + (k - 1);
                rec[(e * nrec_sys + blockIdx.x) * kSmallMax + t] = tmn[t];
            }
            __syncthreads();
        }
    }
}

__global__ void k_extrema(View d, int invert) {
    const Acc A{d.F, d.nbp};
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= d.nb) return;
    if (invert) A.st(F_MIND, g, 1 / A.ld(F_MIND, g));
    else { A.st(F_MIND, g, 0.0); A.st(F_MAXI, g, 0.0); }
}


// ---------------------------------------------------------------------------
// massless test particles in the field of a small system (config of the
// reference: the particle is one more body, after the field bodies, of a
// system with N <= 48, so the single-thread summation order applies)
// ---------------------------------------------------------------------------
// Acceleration of one particle at state (r, v) from the field records T[0..nf)
// (mass.cpp:311-374 for body index nf of an (nf+1)-body system).  The field
// bodies' own sums are unaffected by the particle: its terms sit far below
// their last ulp.  Terms that are exactly zero for a body without spin,
// deformation or receptance (own tidal torque/force, own Lense-Thirring,
// radiation received) are skipped: adding +-0 never changes a sum that
// starts at +0 in round-to-nearest.
__device__ __forceinline__ v3 particle_accel(const JMain *__restrict__ T, int nf, const ddv3 &r, const ddv3 &v,
                                             double sR2, double &mind, double &maxi) {
    double phi = 0;
    v3 nacc{0, 0, 0};
    for (int j = 0; j < nf; ++j) {                          // deform_by, _NK2 class
        const ddv3 rj{dd{T[j].rxh, T[j].rxl}, dd{T[j].ryh, T[j].ryl}, dd{T[j].rzh, T[j].rzl}};
        const v3 d = diff(rj, r);
        const double rr2 = 1 / smax(dot(d, d), sR2 + T[j].sR2);
        const double rr = sqrt(rr2);
        const double tp_dphi = rr * T[j].GM;
        const double tp_dg = rr2 * tp_dphi;
        phi -= tp_dphi;
        nacc += tp_dg * d;
    }
    IMain I;
    I.r = r;
    I.sR2 = sR2;
    I.beta = todbl(v) / kC;
    I.beta2 = dot(I.beta, I.beta);
    I.phi = phi / kC2;
    const v3 naccc = nacc / kC2;
    v3 da{0, 0, 0}, ga{0, 0, 0};
    for (int p = 0; p < nf; ++p) {                          // reactions of earlier bodies' quadrupoles
        const JMain &Tp = T[p];
        bool quad = false;
#pragma unroll
        for (int q = 0; q < 9; ++q) quad |= Tp.C[q] != 0;
        if (!quad) continue;
        const ddv3 rp{dd{Tp.rxh, Tp.rxl}, dd{Tp.ryh, Tp.ryl}, dd{Tp.rzh, Tp.rzl}};
        PairGeo g;
        g.r = diff(r, rp);
        g.rr2 = 1 / smax(dot(g.r, g.r), Tp.sR2 + sR2);
        g.rr = sqrt(g.rr2);
        g.rr3 = g.rr * g.rr2;
        const m3 Cp{v3{Tp.C[0], Tp.C[1], Tp.C[2]}, v3{Tp.C[3], Tp.C[4], Tp.C[5]}, v3{Tp.C[6], Tp.C[7], Tp.C[8]}};
        da -= Tp.GM * ((g.rr3 * g.rr2 * Tp.R2) * quad_field(Cp, g));
    }
    for (int j = 0; j < nf; ++j) {                          // own partner loop
        const JMain &Tj = T[j];
        const PairGeo g = relativity(I, Tj, ga);
        mind = checked_max_into(mind, g.rr);
        maxi = checked_max_into(maxi, g.tp_dg);
        if (Tj.GLx != 0 || Tj.GLy != 0 || Tj.GLz != 0) {    // LENSE_THIRRING, the partner's spin
            const double rcr32 = 2 * Tj.GM / kC * g.rr3;
            const v3 GLj{Tj.GLx, Tj.GLy, Tj.GLz};
            da += cross(rcr32 * (GLj - (3 * dot(GLj, g.r) * g.rr2) * g.r), g.b);
        }
    }
    return (ga + da) + naccc * kC2;                          // gaccel + daccel + naccel
}

// Particle fields (stride pv.npp): r (6), v (6), last acceleration (3),
// min_distance, max_influence; then, per lane of the launch, the derivatives
// of stages kPartLdsStages..24 of the step in progress (3 rows each) and the
// stage sums of stages 1..24 (3 rows each).
enum PField : int { P_R = 0, P_V = 6, P_A = 12, P_MIND = 15, P_MAXI = 16, P_COUNT = 17, P_STAGE = 17 };
// Stage derivatives 0..7 of a step live in LDS (192 B per particle: three
// 256-particle workgroups per CU), 8..24 in the particle's scratch rows;
// every access names its address space, so no access goes through a
// generic pointer.
static_assert(kPartLdsStages == 8, "the LDS share and the occupancy hint of k_particles go together");
typedef __attribute__((address_space(3))) double lds_f64;
typedef __attribute__((address_space(1))) double glb_f64;

// n_steps > 0: advance; n_steps < 0: compute the first acceleration from
// the records; 0: leave the workgroup's particles alone.
// Three wavefronts per SIMD (168 registers; a few values spill to scratch).
__global__ __launch_bounds__(kPartBS) __attribute__((amdgpu_waves_per_eu(3))) void k_particles(PView pv, int n_steps, double dt, DtK dtk,
                                                       const PBlock *__restrict__ blocks) {
    const int p = blockIdx.x * kPartBS + (int)threadIdx.x;
    int64_t entry0 = 0;
    int n_valid = kPartBS;
    if (blocks) {                                           // per-workgroup job offset and length
        entry0 = blocks[blockIdx.x].rec_entry0;
        n_steps = blocks[blockIdx.x].n_steps;
        n_valid = blocks[blockIdx.x].n_valid;
    }
    if (n_steps == 0) return;
    const bool init = n_steps < 0;
    // lanes past the last particle replay the workgroup's first one and store nothing
    const bool valid = p < pv.np && (int)threadIdx.x < n_valid;
    const int q = valid ? p : blockIdx.x * kPartBS;
    double *P = pv.P;
    const int64_t S = pv.npp;
    auto ld = [&](int f) { return P[f * S + q]; };
    ddv3 r{dd{ld(P_R + 0), ld(P_R + 1)}, dd{ld(P_R + 2), ld(P_R + 3)}, dd{ld(P_R + 4), ld(P_R + 5)}};
    ddv3 v{dd{ld(P_V + 0), ld(P_V + 1)}, dd{ld(P_V + 2), ld(P_V + 3)}, dd{ld(P_V + 4), ld(P_V + 5)}};
    double mind = ld(P_MIND), maxi = ld(P_MAXI);
    if (init) {
        const v3 a = particle_accel(pv.rec + entry0 * pv.nrec_sys * kSmallMax, pv.nf, r, v, pv.sR2, mind, maxi);
        if (valid) {
            P[(P_A + 0) * S + q] = a.x; P[(P_A + 1) * S + q] = a.y; P[(P_A + 2) * S + q] = a.z;
            P[P_MIND * S + q] = mind; P[P_MAXI * S + q] = maxi;
        }
        return;
    }
    // stage derivatives a_j (j = 0..24) of this step, per particle
    __shared__ double fa[kPartLdsStages * 3 * kPartBS];
    lds_f64 *fl = (lds_f64 *)fa + threadIdx.x;
    glb_f64 *fg = (glb_f64 *)(P + (int64_t)P_STAGE * S + p);   // this lane's own column (p < npp)
    auto get_l = [&](int j) { return v3{fl[(j * 3 + 0) * kPartBS], fl[(j * 3 + 1) * kPartBS], fl[(j * 3 + 2) * kPartBS]}; };
    auto get_g = [&](int j) {
        const int64_t r = (int64_t)(j - kPartLdsStages) * 3;
        return v3{fg[(r + 0) * S], fg[(r + 1) * S], fg[(r + 2) * S]};
    };
    auto put = [&](int j, v3 a) {
        if (j < kPartLdsStages) {
            fl[(j * 3 + 0) * kPartBS] = a.x; fl[(j * 3 + 1) * kPartBS] = a.y; fl[(j * 3 + 2) * kPartBS] = a.z;
        } else {
            const int64_t r = (int64_t)(j - kPartLdsStages) * 3;
            fg[(r + 0) * S] = a.x; fg[(r + 1) * S] = a.y; fg[(r + 2) * S] = a.z;
        }
    };
    // stage sums sum_l (a_jl) * a_l of stages j = 1..24, kept from stage j
    // (the velocity of stage j is v0 + dt * that sum)
    glb_f64 *fs = fg + (int64_t)kPartStageRows * S;
    auto get_s = [&](int j) {
        const int64_t r = (int64_t)(j - 1) * 3;
        return v3{fs[(r + 0) * S], fs[(r + 1) * S], fs[(r + 2) * S]};
    };
    auto put_s = [&](int j, v3 c) {
        const int64_t r = (int64_t)(j - 1) * 3;
        fs[(r + 0) * S] = c.x; fs[(r + 1) * S] = c.y; fs[(r + 2) * S] = c.z;
    };
    v3 a{ld(P_A + 0), ld(P_A + 1), ld(P_A + 2)};
    for (int step = 0; step < n_steps; ++step) {
        const ddv3 r0 = r, v0 = v;
        put(0, a);
        for (int k = 1; k <= 25; ++k) {
            // x_k = x0 + dt * sum_j a_kj f_j (RungeKutta.impl:148-170); the
            // velocity of stage j is v0 + dt * (sum_l a_jl a_l), that sum as
            // stage j formed it (the same operations, so the same bits)
            ddv3 cv = mkddv(v3{0, 0, 0});
            v3 ca{0, 0, 0};
            const int base = k * (k - 1) / 2;
            const int e1 = c_row0[k + 1];
            int e = c_row0[k];
            for (; e < e1 && c_nzj[e] < kPartLdsStages; ++e) {   // non-zero columns, ascending
                const int j = c_nzj[e];
                const double2 c = c_rk[base + j];
                const dd ckj{c.x, c.y};
                const double fc = c.x + c.y;
                ddv3 vj = v0;
                if (j > 0) vj = add(v0, mkddv(dt * get_s(j)));
                cv = add(cv, scale(ckj, vj));
                ca += fc * get_l(j);
            }
            for (; e < e1; ++e) {
                const int j = c_nzj[e];
                const double2 c = c_rk[base + j];
                const dd ckj{c.x, c.y};
                const double fc = c.x + c.y;
                const ddv3 vj = add(v0, mkddv(dt * get_s(j)));
                cv = add(cv, scale(ckj, vj));
                ca += fc * get_g(j);
            }
            if (k < 25) put_s(k, ca);
            r = add(r0, scale(mkdd(dt), cv));
            v = add(v0, mkddv(dt * ca));
            const JMain *T = pv.rec + ((entry0 + 25 * (int64_t)step + (k - 1)) * pv.nrec_sys) * kSmallMax;
            a = particle_accel(T, pv.nf, r, v, pv.sR2, mind, maxi);
            if (k < 25) put(k, a);
        }
    }
    if (!valid) return;
    const double out[P_COUNT] = {r.x.hi, r.x.lo, r.y.hi, r.y.lo, r.z.hi, r.z.lo,
                                 v.x.hi, v.x.lo, v.y.hi, v.y.lo, v.z.hi, v.z.lo, a.x, a.y, a.z, mind, maxi};
#pragma unroll
    for (int f = 0; f < P_COUNT; ++f) P[f * S + q] = out[f];
}

void launch_particles(const PView &pv, int n_steps, double dt, const DtK &dtk, hipStream_t st, const PBlock *blocks) {
    if (pv.np) hipLaunchKernelGGL(k_particles, dim3((pv.np + kPartBS - 1) / kPartBS), dim3(kPartBS), 0, st, pv,
                                  n_steps, dt, dtk, blocks);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void set_rk_table(const double *tab650, hipStream_t st) {
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_rk), tab650, sizeof(double) * 650, 0, hipMemcpyHostToDevice, st);
    static int nzj[325], row0[27];
    int e = 0;
    for (int k = 0; k <= 25; ++k) {
        row0[k] = e;
        if (k == 0) continue;
        for (int j = 0; j < k; ++j)
            if (tab650[2 * (k * (k - 1) / 2 + j)] != 0) nzj[e++] = j;
    }
    row0[26] = e;
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_nzj), nzj, sizeof(nzj), 0, hipMemcpyHostToDevice, st);
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_row0), row0, sizeof(row0), 0, hipMemcpyHostToDevice, st);
}

void launch_deform_mixed(const View &d, const int2 *work, int nwork, const int4 *blk16, int nblk, int prepare,
                         int np_rows, int max_n, hipStream_t st) {
    const NpTail np{nwork + nblk, np_rows, (max_n + 63) / 64};
    const int grid = nwork + nblk + np_rows * np.chunks;
    if (grid) hipLaunchKernelGGL(k_deform_mixed, dim3(grid), dim3(64), 0, st, d, work, nwork, blk16, prepare, np);
}
void launch_deform(const View &d, const int4 *blk_light, int n_light, const int4 *blk_damped, int n_damped,
                   int prepare, hipStream_t st) {
    if (n_damped) hipLaunchKernelGGL(k_deform_damped<kBS>, dim3(n_damped), dim3(kBS), 0, st, d, blk_damped, prepare);
    if (n_light) hipLaunchKernelGGL(k_deform_light<kBS>, dim3(n_light), dim3(kBS), 0, st, d, blk_light, prepare);
}
void launch_main_pool(const View &d, const int2 *blk, int nblk, hipStream_t st) {
    if (nblk) hipLaunchKernelGGL(k_main_pool<kBS>, dim3(nblk), dim3(kBS), 0, st, d, blk);
}
void launch_main_k4_items(const View &d, const int2 *blk16, int nblk, int nrows, int maxT, hipStream_t st) {
    const int grid = nblk + nrows * maxT;
    if (grid) hipLaunchKernelGGL(k_main_k4_items, dim3(grid), dim3(64), 0, st, d, blk16, nblk, nrows, maxT);
}
void launch_main_k4(const View &d, const int2 *blk16, int nblk, hipStream_t st) {
    if (nblk) hipLaunchKernelGGL(k_main_k4, dim3(nblk), dim3(64), 0, st, d, blk16);
}
// Four wavefronts per body while every body's workgroup is resident at once
// (52 KB of LDS each: three per CU): the pair terms of a 504-body system are
// formed in two rounds instead of four (38 us instead of 44 us per launch on
// one MI355X); the ordered sums are the same.
void launch_main_wpb(const View &d, const int2 *work, int nwork, int max_n, bool fused_np, int waves, hipStream_t st) {
    if (!nwork) return;
    int W = waves;
    if (W == 0) W = (max_n > 128 && nwork <= 768) ? 4 : (max_n > 64 && nwork <= 1536) ? 2 : 1;
    if (fused_np && W == 1) W = 2;
    if (fused_np && W == 4) hipLaunchKernelGGL((k_main_wpb<true, 4>), dim3(nwork), dim3(256), 0, st, d, work);
    else if (fused_np) hipLaunchKernelGGL((k_main_wpb<true, 2>), dim3(nwork), dim3(128), 0, st, d, work);
    else if (W == 4) hipLaunchKernelGGL((k_main_wpb<false, 4>), dim3(nwork), dim3(256), 0, st, d, work);
    else if (W == 2) hipLaunchKernelGGL((k_main_wpb<false, 2>), dim3(nwork), dim3(128), 0, st, d, work);
    else hipLaunchKernelGGL((k_main_wpb<false, 1>), dim3(nwork), dim3(64), 0, st, d, work);
}
// Wavefronts per body: several while every body's workgroup still fits on
// the GPU at once (the ordered sums are serial, so extra wavefronts only pay
// while CUs would otherwise be idle; a large batch is faster with one
// wavefront per body).  Four, not eight, for a 504-body system: two 256-lane
// workgroups then share each CU, one round of workgroups instead of two
// (31 us instead of 43-58 us per launch on one MI355X).
void launch_deform_wpb(const View &d, const int2 *work, int nwork, int max_n, int prepare, int np_rows, int waves,
                       hipStream_t st) {
    if (!nwork && !np_rows) return;
    const int chunks = (max_n + 63) / 64;
    int W = waves;
    if (W == 0) W = (chunks >= 4 && nwork <= 1280) ? 4 : (chunks >= 2 && nwork <= 2816) ? 2 : 1;
    const NpTail np{nwork, np_rows, (max_n + 63) / 64};
    const dim3 grid(nwork + (np_rows * np.chunks + W - 1) / W), block(64 * W);
    if (np_rows) {
        if (W == 4) hipLaunchKernelGGL((k_deform_wpb<4, true>), grid, block, 0, st, d, work, prepare, np);
        else if (W == 2) hipLaunchKernelGGL((k_deform_wpb<2, true>), grid, block, 0, st, d, work, prepare, np);
        else hipLaunchKernelGGL((k_deform_wpb<1, true>), grid, block, 0, st, d, work, prepare, np);
    } else {
        if (W == 4) hipLaunchKernelGGL((k_deform_wpb<4, false>), grid, block, 0, st, d, work, prepare, np);
        else if (W == 2) hipLaunchKernelGGL((k_deform_wpb<2, false>), grid, block, 0, st, d, work, prepare, np);
        else hipLaunchKernelGGL((k_deform_wpb<1, false>), grid, block, 0, st, d, work, prepare, np);
    }
}
void launch_main_serial(const View &d, const int2 *blk, int nblk, hipStream_t st) {
    if (nblk) hipLaunchKernelGGL(k_main_serial<kBS>, dim3(nblk), dim3(kBS), 0, st, d, blk);
}
void launch_np_field(const View &d, int nrows, int max_n, hipStream_t st) {
    if (nrows) hipLaunchKernelGGL(k_np_field, dim3((max_n + 63) / 64, nrows), dim3(64), 0, st, d, nrows);
}
void launch_np(const View &d, int nrows, int max_n, int maxT, const int2 *recv_blk, int nrecv, bool with_field,
               hipStream_t st, bool with_item) {
    if (!nrows) return;
    if (with_field) hipLaunchKernelGGL(k_np_field, dim3((max_n + 63) / 64, nrows), dim3(64), 0, st, d, nrows);
    if (with_item) hipLaunchKernelGGL(k_np_item, dim3(nrows * maxT), dim3(64), 0, st, d, nrows, maxT);
    if (nrecv) hipLaunchKernelGGL(k_np_recv, dim3(nrecv), dim3(kBS), 0, st, d, recv_blk, maxT);
}
void launch_stage(const View &d, int k, double dt, double dt_k, double tadd, const int64_t *istep, hipStream_t st) {
    hipLaunchKernelGGL(k_stage, dim3(2 * ((d.nb + 63) / 64)), dim3(192), 0, st, d, k, dt, dt_k, tadd, istep);
}
void launch_step_count(int64_t *istep, hipStream_t st) {
    hipLaunchKernelGGL(k_step_count, dim3(1), dim3(64), 0, st, istep);
}
// 128 threads for systems of at most 16 bodies (their pairs fit one round),
// 256 above (fewer rounds of pairs; the extra waves cost the small ones time
// at every barrier)
void launch_small(const View &d, const int *sys_ids, int nsys, int max_n, int n_steps, double dt, const DtK &dtk,
                  int64_t i_step0, JMain *rec, hipStream_t st) {
    if (!nsys) return;
    if (max_n <= 16)
        hipLaunchKernelGGL(k_small<2 * kSmallMax>, dim3(nsys), dim3(2 * kSmallMax), 0, st, d, sys_ids, n_steps, dt, dtk,
                           i_step0, rec, nsys);
    else
        hipLaunchKernelGGL(k_small<4 * kSmallMax>, dim3(nsys), dim3(4 * kSmallMax), 0, st, d, sys_ids, n_steps, dt, dtk,
                           i_step0, rec, nsys);
}
void launch_extrema(const View &d, int invert, hipStream_t st) {
    hipLaunchKernelGGL(k_extrema, dim3((d.nb + 255) / 256), dim3(256), 0, st, d, invert);
}

}  // namespace eph
