// Device data layout: one field-major array of doubles per batch.
//
// Field f of global body g lives at F[f * nbp + g] (nbp = bodies rounded up
// to 64), so consecutive threads touch consecutive addresses.  Double-double
// quantities are split into hi and lo fields.  Bodies of system s occupy
// [sys_off[s], sys_off[s+1]).
#pragma once
#include <stdint.h>

namespace eph {

enum Field : int {
    // evolving state
    F_R = 0,            // 6: x.hi x.lo y.hi y.lo z.hi z.lo
    F_V = 6,            // 6
    F_GL = 12,          // 6
    F_S = 18,           // 18: axis x (x,y,z hi/lo), axis y, axis z
    F_W = 36,           // 3
    F_GM = 39,
    F_EXJ2 = 40,
    // constants
    F_GM0 = 41, F_DGM, F_DJ2, F_RAD, F_K2, F_K2R, F_TD, F_TDF, F_LUM, F_RR2G4C,
    F_A, F_R2, F_SR2,
    F_CST = 54,         // 9 C_static
    F_INERTIA = 63, F_RECPT = 64,
    // accel outputs
    F_PHI = 65,         // final phi (m^2/s^2)
    F_PHIC = 66,        // phi / c^2 (input of the 1PN term)
    F_BETA2 = 67,
    F_NACC = 68,        // 3 final Newtonian acceleration
    F_NACCC = 71,       // 3 naccel / c^2
    F_DACC = 74,        // 3
    F_DTQ = 77,         // 3
    F_GACC = 80,        // 3
    F_BETA = 83,        // 3
    F_GI = 86,          // 9
    F_CPOT = 95,        // 9
    F_EGRAD = 104,      // 3
    F_IDACC = 107, F_IDTQ = 110, F_JDACC = 113, F_JDTQ = 116,
    F_MIND = 119, F_MAXI = 120,
    // RK12 step-start state
    F_X0R = 121,        // 6
    F_X0V = 127,        // 6
    F_X0GL = 133,       // 6
    F_X0S = 139,        // 18
    F_X0W = 157,        // 3
    F_STAGE = 160,      // 25 stages x 15: v(6 dd), a(3), w(3), dtorque(3)
    F_COUNT = 160 + 25 * 15
};
constexpr int kStageFields = 15;

// Non-point force items of one system (geopotential or ring of one body).
struct NPItem {
    int32_t body;      // global body index
    int32_t kind;      // 0 = geopotential, 1 = ring
    int32_t model;     // index into the geopotential or ring table
    int32_t chain;     // chain index in the pooled non-point rounds
    int32_t sys;       // system
    int32_t an_off;    // offset (in bodies) of this item's row of fields
};

struct SysMeta {
    int32_t off, n;        // body range
    int32_t item0, nitem;  // non-point items [item0, item0+nitem), creation order
    int32_t T;             // pooled non-point rounds (0 = none)
    int32_t serial;        // 1: single-thread summation order
};

}  // namespace eph
