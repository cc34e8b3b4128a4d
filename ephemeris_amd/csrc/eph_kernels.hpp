// Kernel-side view of a resident batch and the launch entry points.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/ephemeris_hip.h"
#include "eph_layout.hpp"
#include "eph_math.hpp"
#include "eph_terms.hpp"

namespace eph {

constexpr int kBS = 64;          // bodies per workgroup in the j-loop kernels (one wave64)
constexpr int kSmallMax = 64;    // largest system integrated by one workgroup

// sum of the non-zero tableau weights of each stage (dt_k), index 1..25
struct DtK { double v[26]; };

constexpr int kPartBS = 256;     // particles per workgroup
// A particle's stage derivatives 0..kPartLdsStages-1 of the step in progress
// live in LDS (24 B each per particle), the rest and the stage sums of
// stages 1..24 in per-particle scratch rows after the particle's state:
// kPartScratchRows rows, which every particle pool allocates.
constexpr int kPartLdsStages = 8;
constexpr int kPartStageRows = 3 * (25 - kPartLdsStages);        // derivatives kPartLdsStages..24
constexpr int kPartScratchRows = kPartStageRows + 3 * 24;          // and the stage sums 1..24

// test particles in the recorded field of one small system
struct PView {
    double *P;                   // particle fields, field-major, stride npp
    int64_t npp;
    int np;
    const JMain *rec;            // field records of the launch (see k_small)
    int nrec_sys;                // systems interleaved in rec (the field is the first)
    int nf;                      // field bodies
    double sR2;                  // soft radius^2 of the particles (0 for R > 0)
};

struct View {
    double *F;                   // field-major state, F[f * nbp + g]
    int64_t nbp;                 // field stride
    int nb;                      // bodies in the batch
    const SysMeta *sys;
    const NPItem *items;
    const eph_geopotential *gps;
    const eph_ring *rings;
    const int32_t *np_gp;        // per body: geopotential model or -1
    const int32_t *np_ring;      // per body: ring model or -1
    const int32_t *bsys;         // per body: system index
    const int32_t *perm;         // per system slot: body, bodies with delayed tides first
    double *an;                  // non-point field per (item, partner), 3 doubles
    double *part;                // per (item, round) partial sums, 6 doubles
    const double *t0;            // per system: t_eph (hi+lo) at the start of integrate
    // single-thread-order systems: the field of each body's geopotential and
    // ring at every partner of its system, formed before the main force
    // (rows of 3 doubles per partner; -1: the body has no such model), or null
    const int32_t *srow_gp, *srow_ring;
    const double *san;
};

void set_rk_table(const double *tab650, hipStream_t st);
void launch_deform(const View &d, const int4 *blk_light, int n_light, const int4 *blk_damped, int n_damped,
                   int prepare, hipStream_t st);
void launch_main_pool(const View &d, const int2 *blk, int nblk, hipStream_t st);
// constant-time-lag bodies (work = {system, body}, nwork a multiple of 8),
// one wavefront each, and the other bodies four lanes each (blk16 =
// (system, first slot, count <= 16, 0) in the class-sorted slots) in one
// launch, with the item fields
void launch_deform_mixed(const View &d, const int2 *work, int nwork, const int4 *blk16, int nblk, int prepare,
                         int np_rows, int max_n, hipStream_t st);
// four lanes per body; blk16 = (system, first body) of every 16 bodies
void launch_main_k4(const View &d, const int2 *blk16, int nblk, hipStream_t st);
// the same with the items' partial sums of launch_np (rows, rounds) in extra
// workgroups of the launch
void launch_main_k4_items(const View &d, const int2 *blk16, int nblk, int nrows, int maxT, hipStream_t st);
void launch_main_serial(const View &d, const int2 *blk, int nblk, hipStream_t st);
// latency path, one workgroup of 1-4 wavefronts per body; work = {system,
// body index in system}
// fused_np: the receiver and own-item sums of the non-point pass run inside
// each body's workgroup, after the item fields are formed (launch_np_field or
// the extra workgroups of a deform launch; every thread-pool-order system
// has at most 512 bodies and 64 non-point items)
// waves: wavefronts per body, 0 = chosen from the batch's shape, else 1, 2 or
// 4 (the fused kernel has at least 2)
void launch_main_wpb(const View &d, const int2 *work, int nwork, int max_n, bool fused_np, int waves, hipStream_t st);
void launch_np_field(const View &d, int nrows, int max_n, hipStream_t st);
constexpr int kNpFusedMaxN_h = 512, kNpFusedMaxItems_h = 64;
// one workgroup of up to 4 wavefronts per body (partner chunks evaluated in
// parallel, summed in order); max_n = largest system of the batch
// np_rows > 0: the launch also evaluates the non-point item fields
// (launch_np_field) in extra workgroups
// (waves as for launch_main_wpb)
void launch_deform_wpb(const View &d, const int2 *work, int nwork, int max_n, int prepare, int np_rows, int waves,
                       hipStream_t st);
// with_item = false: the items' partial sums were formed already
// (launch_main_k4_items)
void launch_np(const View &d, int nrows, int max_n, int maxT, const int2 *recv_blk, int nrecv, bool with_field,
               hipStream_t st, bool with_item = true);
// istep: optional device step counter (graph replays); launch_step_count adds 1
void launch_stage(const View &d, int k, double dt, double dt_k, double tadd, const int64_t *istep, hipStream_t st);
void launch_step_count(int64_t *istep, hipStream_t st);
void launch_extrema(const View &d, int invert, hipStream_t st);
// per-workgroup schedule of a particle launch: first record entry (25 per
// step), number of steps (< 0: first acceleration from entry rec_entry0,
// 0: idle) and number of particles of the workgroup (>= 1)
struct PBlock { int64_t rec_entry0; int32_t n_steps, n_valid; };
// n_steps < 0 computes the first acceleration from record entry 0
void launch_particles(const PView &pv, int n_steps, double dt, const DtK &dtk, hipStream_t st,
                      const PBlock *blocks = nullptr);
void launch_small(const View &d, const int *sys_ids, int nsys, int max_n, int n_steps, double dt, const DtK &dtk,
                  int64_t i_step0, JMain *rec, hipStream_t st);

}  // namespace eph
