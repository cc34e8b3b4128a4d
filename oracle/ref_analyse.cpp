/*
 * ORACLE / golden-vector harness -- test infrastructure only, built into
 * oracle/_ref/ by `make ref` when the reference checkout is present.
 *
 * Compiles the reference's own src/physics/barycen_analyzer.cpp
 * (msystem::analyse, bsystem::fill_tid) and runs it on bodies given as
 * o_mass.  msystem keeps its bodies private; it befriends a class named
 * ephemeris_generator (src/physics/mass.h:198-200), which this file defines
 * to fill the bodies and read the resulting hierarchy.  The msystem object is
 * never destroyed: its destructor is defined in mass_loader.cpp, which is
 * not compiled here.
 */
#include "physics/barycen_analyzer.cpp"
#include "eph_oracle.h"

#include <new>

#define EXPORT extern "C" __attribute__((visibility("default")))

class ephemeris_generator {
public:
    static int64_t analyse(const o_mass *ms, int64_t n, double t_hi, double t_lo, int64_t *ids, int64_t cap,
                           int64_t *children, int64_t child_cap, double *gm_sys) {
        msystem *sys = new msystem();
        sys->mlist.resize(n);
        for (int64_t i = 0; i < n; ++i) {
            mass &m = sys->mlist[i];
            const o_mass &o = ms[i];
            m.r = mpvec(real(o.r.x.hi, o.r.x.lo), real(o.r.y.hi, o.r.y.lo), real(o.r.z.hi, o.r.z.lo));
            m.v = mpvec(real(o.v.x.hi, o.v.x.lo), real(o.v.y.hi, o.v.y.lo), real(o.v.z.hi, o.v.z.lo));
            m.GM = o.GM; m.R = o.R; m.sR2 = o.sR2; m.sid = o.sid;
        }
        sys->t_eph = real(t_hi, t_lo);
        sys->analyse(true);
        const bsystem &bl = sys->blist;
        int64_t nc = 0;
        for (int64_t i = 0; i < (int64_t)bl.size() && i < cap; ++i) {
            const barycen &b = bl[i];
            int64_t *o = ids + 7 * i;
            o[0] = b.pid; o[1] = b.hid; o[2] = b.gid; o[3] = b.tid; o[4] = b.mid;
            o[5] = (int64_t)b.children.size();
            o[6] = nc;
            gm_sys[2 * i] = b.GM_sys.hi;
            gm_sys[2 * i + 1] = b.GM_sys.lo;
            for (auto c : b.children) {
                if (nc < child_cap) children[nc] = c;
                ++nc;
            }
        }
        return (int64_t)bl.size();
    }
};

/* msystem::analyse(true) on n bodies at time t; writes per barycen
 * {pid, hid, gid, tid, mid, number of children, offset of its children} and
 * GM_sys {hi, lo}; returns the number of barycens */
EXPORT int64_t ref_analyse(const o_mass *ms, int64_t n, double t_hi, double t_lo, int64_t *ids, int64_t cap,
                           int64_t *children, int64_t child_cap, double *gm_sys) {
    return ephemeris_generator::analyse(ms, n, t_hi, t_lo, ids, cap, children, child_cap, gm_sys);
}
