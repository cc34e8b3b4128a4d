// Orbit-structure analysis (barycen hierarchy), the combined (hierarchical)
// integrator and the binary checkpoint of MSystem.  Host code only: the
// combined integrator runs the full system's long step through the GPU or
// CPU back end and the subsystems' short steps on the CPU back end, as the
// reference does.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <thread>

#include "eph_cpu.hpp"
#include "eph_msystem.hpp"
#include "eph_runtime.hpp"

namespace eph {

namespace {

v3 V(const eph_vec &a) { return v3{a.x, a.y, a.z}; }
eph_vec E(v3 a) { return eph_vec{a.x, a.y, a.z}; }
m3 M(const eph_mat &a) { return m3{V(a.x), V(a.y), V(a.z)}; }
eph_mat E(const m3 &a) { return eph_mat{E(a.x), E(a.y), E(a.z)}; }
ddv3 DV(const eph_ddvec &a) { return ddv3{dd{a.x.hi, a.x.lo}, dd{a.y.hi, a.y.lo}, dd{a.z.hi, a.z.lo}}; }
eph_ddvec E(const ddv3 &a) { return eph_ddvec{{a.x.hi, a.x.lo}, {a.y.hi, a.y.lo}, {a.z.hi, a.z.lo}}; }
ddm3 DM(const eph_ddmat &a) { return ddm3{DV(a.x), DV(a.y), DV(a.z)}; }
eph_ddmat E(const ddm3 &a) { return eph_ddmat{E(a.x), E(a.y), E(a.z)}; }

const ddv3 kZero3{dd{0, 0}, dd{0, 0}, dd{0, 0}};
constexpr int64_t kMaxAnalysePasses = 64;
// real(x) != real(x): double-double comparison of dfloat_t.h:42-47, NaN test
bool dd_isnan(dd a) { return todbl(sub(a, a)) != 0 || a.hi != a.hi; }
bool dd_eq(dd a, dd b) { return todbl(sub(a, b)) == 0; }

}  // namespace

// ---------------------------------------------------------------------------
// msystem::analyse (src/physics/barycen_analyzer.cpp:12-310) and
// bsystem::fill_tid (:312-333)
// ---------------------------------------------------------------------------
dd MSystem::analyse(bool reconstruct) {
    struct BData { double minr = 0, factor = 0, maxinfl = 0; bool del = false; };
    int64_t old_bn = reconstruct ? 0 : (int64_t)blist.size();
    const int64_t mn = size();
    if (old_bn && dd_eq(t_barycen, t_eph_)) return t_update;
    t_barycen = t_eph_;
    std::vector<Barycen> bl;
    std::vector<BData> dl;
    if (old_bn) {
        bl.resize((size_t)old_bn);
        dl.resize((size_t)old_bn);
        for (int64_t i = 0; i < old_bn; ++i) {
            Barycen &bi = bl[(size_t)i];
            bi.pid = blist[(size_t)i].pid;
            bi.hid = blist[(size_t)i].hid;
            bi.gid = blist[(size_t)i].gid;
            bi.mid = bi.hid < 0 ? blist[(size_t)i].mid : -1;
        }
    } else {
        bl.resize((size_t)mn);
        dl.resize((size_t)mn);
        for (int64_t i = 0; i < mn; ++i) bl[(size_t)i].mid = i;
    }

    const double cbrt_half = std::cbrt(.5);
    bool diff = false;
    int64_t passes = 0;
    analysis_settled = true;
    do {
        if (diff) old_bn = 0;
        const int64_t bn = (int64_t)bl.size();
        // The reference repeats this until nothing changes; on some inputs
        // (three bodies whose mutual influences tie) it alternates between two
        // hierarchies for ever and its loader never returns.  Here the
        // analysis gives up after a bounded number of passes and leaves the
        // previous structure in place.
        if (++passes > kMaxAnalysePasses) {
            analysis_settled = false;
            return t_update;
        }
        // state of the barycens: r, v, GM and the systems' r_sys, v_sys, GM_sys
        for (int64_t i = 0; i < bn; ++i)
            if (!dl[(size_t)i].del) {
                bl[(size_t)i].GM_sys = dd{0, 0};
                bl[(size_t)i].r_sys = kZero3;
                bl[(size_t)i].v_sys = kZero3;
            }
        for (int64_t i = 0; i < bn; ++i) {
            Barycen &bi = bl[(size_t)i];
            if (bi.mid < 0) continue;
            const eph_mass &mi = mlist[(size_t)bi.mid];
            const dd mGM = mkdd(mi.GM);
            const ddv3 rGM = eph::scale(mGM, DV(mi.r)), vGM = eph::scale(mGM, DV(mi.v));
            bi.r = DV(mi.r);
            bi.v = DV(mi.v);
            bi.GM = mGM;
            bi.r_sys = add(bi.r_sys, rGM);
            bi.v_sys = add(bi.v_sys, vGM);
            bi.GM_sys = add(bi.GM_sys, mGM);
            for (int64_t bp = bi.pid; bp >= 0 && !dl[(size_t)bp].del; bp = bl[(size_t)bp].pid) {
                Barycen &p = bl[(size_t)bp];
                p.GM_sys = add(p.GM_sys, mGM);
                p.r_sys = add(p.r_sys, rGM);
                p.v_sys = add(p.v_sys, vGM);
            }
        }
        for (int64_t i = 0; i < bn; ++i) {
            Barycen &bi = bl[(size_t)i];
            if (bi.mid >= 0 || dl[(size_t)i].del) continue;
            const Barycen &h = bl[(size_t)bi.hid], &g = bl[(size_t)bi.gid];
            bi.GM = add(h.GM_sys, g.GM_sys);
            bi.r = divide(add(h.r_sys, g.r_sys), bi.GM);
            bi.v = divide(add(h.v_sys, g.v_sys), bi.GM);
        }
        for (int64_t i = 0; i < bn; ++i)
            if (!dl[(size_t)i].del) {
                Barycen &bi = bl[(size_t)i];
                bi.r_sys = divide(bi.r_sys, bi.GM_sys);
                bi.v_sys = divide(bi.v_sys, bi.GM_sys);
            }
        // minimum separation and influence factor of each binary
        for (int64_t i = 0; i < bn; ++i) {
            if (dl[(size_t)i].del) continue;
            Barycen &bi = bl[(size_t)i];
            BData &di = dl[(size_t)i];
            if (bi.mid >= 0) { di.minr = 0; continue; }
            const v3 r = todbl(sub(bl[(size_t)bi.gid].r_sys, bl[(size_t)bi.hid].r_sys));
            di.minr = length(r);
            const double gmh = todbl(bl[(size_t)bi.hid].GM_sys), gmg = todbl(bl[(size_t)bi.gid].GM_sys);
            double q = gmh > gmg ? gmg / gmh : gmh / gmg;
            di.minr /= 1 - 1.0001 * (std::cbrt(q) - q);
            q += 1;
            q *= q;
            di.factor = 1 / q;
        }
        int64_t n_root = 0, root = -1;
        // influence of the present parent
        for (int64_t i = 0; i < bn; ++i) {
            if (dl[(size_t)i].del) continue;
            Barycen &bi = bl[(size_t)i];
            BData &di = dl[(size_t)i];
            if (bi.pid < 0 || dl[(size_t)bi.pid].del) {
                bi.pid = -1;
                di.maxinfl = 0;
                ++n_root;
                root = i;
            } else {
                const Barycen &p = bl[(size_t)bi.pid];
                if (p.hid == i || p.gid == i) {                 // one of a binary
                    const double r1 = dl[(size_t)bi.pid].minr, r2 = r1 * r1;
                    double rmax = std::max(dl[(size_t)p.hid].minr, dl[(size_t)p.gid].minr);
                    if (old_bn) rmax *= cbrt_half;
                    if (r1 < rmax) {
                        bi.pid = -1;
                        di.maxinfl = 0;
                    } else {
                        di.maxinfl = todbl(p.GM) / (r2 * r1);
                        di.maxinfl *= 2;
                    }
                } else {
                    const v3 r = todbl(sub(bi.r_sys, p.r));
                    const double r2 = dot(r, r), r1 = sqrt(r2);
                    double rmax = std::max(dl[(size_t)bi.pid].minr, di.minr);
                    if (old_bn) rmax *= cbrt_half;
                    if (r1 < rmax) {
                        bi.pid = -1;
                        di.maxinfl = 0;
                    } else
                        di.maxinfl = todbl(p.GM) / (r2 * r1);
                }
            }
            if (old_bn) di.maxinfl *= 2;
        }
        if (n_root != 1) root = -1;
        diff = false;
        // the most influential parent of each node
        for (int64_t i = 0; i < bn; ++i) {
            if (i == root) continue;
            Barycen &bi = bl[(size_t)i];
            BData &di = dl[(size_t)i];
            if (di.del) continue;
            const int64_t pid = bi.pid;
            int64_t cid = -1;
            if (pid >= 0 && bl[(size_t)pid].mid < 0) {
                if (bl[(size_t)pid].hid == i) cid = bl[(size_t)pid].gid;
                if (bl[(size_t)pid].gid == i) cid = bl[(size_t)pid].hid;
            }
            const double soft_r2 = bi.mid >= 0 ? mlist[(size_t)bi.mid].sR2 : 0;
            for (int64_t j = 0; j < bn; ++j) {
                if (i == j || pid == j || cid == j) continue;
                const Barycen &bj = bl[(size_t)j];
                const BData &dj = dl[(size_t)j];
                if (dj.del) continue;
                if (bj.mid >= 0 && mlist[(size_t)bj.mid].R < 0) continue;    // soft masses parent nothing
                const v3 r = todbl(sub(bj.r, bi.r));
                const double r2 = dot(r, r), r1 = sqrt(r2);
                if (!(r2 > soft_r2)) continue;
                if (!(r1 > di.minr && r1 > dj.minr)) continue;
                double infl = todbl(bj.GM) / (r2 * r1);
                int64_t tj = j, tp = bj.pid;
                while (tp >= 0) {                                // far from a binary: reduced influence
                    const Barycen &bp = bl[(size_t)tp];
                    if (!(bp.mid < 0 && (bp.hid == tj || bp.gid == tj))) break;
                    const v3 dr = todbl(sub(bp.r_sys, bi.r));
                    const double minr = dl[(size_t)tp].minr;
                    if (dot(dr, dr) > minr * minr) infl *= dl[(size_t)tp].factor;
                    tj = tp;
                    tp = bp.pid;
                }
                if (infl > di.maxinfl) {
                    di.maxinfl = infl;
                    bi.pid = j;
                    diff = true;
                }
            }
        }
        // delete unbound barycens
        for (int64_t i = mn; i < bn; ++i) {
            Barycen &bi = bl[(size_t)i];
            if (dl[(size_t)i].del) continue;
            if (bl[(size_t)bi.hid].pid != i || bl[(size_t)bi.gid].pid != i) {
                dl[(size_t)i].del = true;
                diff = true;
                int64_t ti = i, tp = bi.pid;
                while (tp >= 0 && (bl[(size_t)tp].hid == ti || bl[(size_t)tp].gid == ti)) {
                    dl[(size_t)tp].del = true;
                    ti = tp;
                    tp = bl[(size_t)ti].pid;
                }
            }
        }
        // create new barycens of mutual parents
        for (int64_t i = 0; i < bn; ++i) {
            if (dl[(size_t)i].del) continue;
            const int64_t j = bl[(size_t)i].pid;
            if (j < 0 || dl[(size_t)j].del || j >= bn) continue;
            if (bl[(size_t)j].pid != i) {
                if (todbl(bl[(size_t)j].GM_sys) < todbl(bl[(size_t)i].GM_sys)) {
                    bl[(size_t)i].pid = -1;
                    diff = true;
                }
                continue;
            }
            if (j < i) continue;
            const int64_t k = (int64_t)bl.size();
            bl[(size_t)i].pid = k;
            bl[(size_t)j].pid = k;
            Barycen bk;
            bk.pid = old_bn && blist[(size_t)i].pid == blist[(size_t)j].pid ? blist[(size_t)i].pid : -1;
            bk.hid = i;
            bk.gid = j;
            bk.mid = -1;
            bl.push_back(bk);
            dl.push_back(BData{});
            diff = true;
        }
    } while (diff);

    if (old_bn) {
        if (dd_isnan(t_update)) {
            for (int64_t i = 0; i < old_bn; ++i) {
                const Barycen &src = bl[(size_t)i];
                Barycen &dst = blist[(size_t)i];
                dst.r = src.r;
                dst.v = src.v;
                dst.GM = src.GM;
                dst.r_sys = src.r_sys;
                dst.v_sys = src.v_sys;
                dst.GM_sys = src.GM_sys;
            }
            t_update = t_eph_;
        }
        return t_update;
    }

    // remove deleted barycens
    std::vector<int64_t> newindex(bl.size() + 1, -1);
    int64_t bsize = 0;
    for (int64_t i = 0, n = (int64_t)bl.size(); i < n; ++i) {
        if (dl[(size_t)i].del) continue;
        newindex[(size_t)i + 1] = bsize;
        if (i != bsize) bl[(size_t)bsize] = bl[(size_t)i];
        ++bsize;
    }
    bl.resize((size_t)bsize);
    for (auto &bi : bl) {
        bi.pid = newindex[(size_t)(bi.pid + 1)];
        bi.hid = newindex[(size_t)(bi.hid + 1)];
        bi.gid = newindex[(size_t)(bi.gid + 1)];
    }
    // host is the heavier of a binary; children lists
    for (int64_t i = 0; i < bsize; ++i) {
        Barycen &bi = bl[(size_t)i];
        if (bi.pid >= 0) {
            Barycen &p = bl[(size_t)bi.pid];
            if (!(p.hid == i || p.gid == i)) p.children.push_back(i);
        }
        if (bi.hid >= 0 && todbl(bl[(size_t)bi.hid].GM_sys) < todbl(bl[(size_t)bi.gid].GM_sys))
            std::swap(bi.hid, bi.gid);
    }
    // fill tid and mid (bsystem::fill_tid)
    for (int64_t i = 0; i < bsize; ++i) {
        Barycen &bi = bl[(size_t)i];
        int64_t oid = i, pid = bi.pid;
        while (pid >= 0 && bl[(size_t)pid].hid == oid) {
            oid = pid;
            pid = bl[(size_t)pid].pid;
        }
        bi.tid = oid;
        if (bi.mid < 0) {
            int64_t hid = i;
            while (bl[(size_t)hid].hid >= 0) hid = bl[(size_t)hid].hid;
            bi.mid = bl[(size_t)hid].mid;
        }
    }
    blist.swap(bl);
    t_update = t_eph_;
    return t_update;
}

// ---------------------------------------------------------------------------
// msystem::combined_integrate (src/integrators/mass_combined.cpp:40-406)
// ---------------------------------------------------------------------------
namespace {

constexpr double kVeryBig = 1.e16;          // VERY_BIG_VALUE, mass_combined.cpp:9
constexpr double kRotationStepLimit = 10;   // ROTATION_STEP_LIMIT, mass_combined.cpp:11

void mass_copy(eph_mass &dst, const eph_mass &src) {   // mass_combined.cpp:15-21
    dst.r = src.r;
    dst.v = src.v;
    dst.GL = src.GL;
    dst.s = src.s;
    dst.w = src.w;
}

// mass::scale (mass.cpp:8-26) of a body copy
void scale_mass(eph_mass &m, double factor) {
    const double f2 = factor * factor, f3 = factor * f2;
    m.R *= factor;
    m.R2 *= f2;
    m.sR2 *= f2;
    m.GL = E(eph::scale(mkdd(f2), DV(m.GL)));
    m.GI = E(f2 * M(m.GI));
    m.GM *= f3;
    m.GM0 *= f3;
    m.dGM *= f3;
    m.lum *= f2;
    m.recpt *= factor;
    m.rR2G_4c *= f3;
}

// bodies and their models of one subsystem, as msystem::push_back builds them
struct SubBuilder {
    std::vector<eph_mass> m;
    std::vector<eph_geopotential> gps;
    std::vector<eph_ring> rgs;
    void push(eph_mass b, const MSystem &from) {
        if (b.gp >= 0) { gps.push_back(from.geopotentials()[(size_t)b.gp]); b.gp = (int64_t)gps.size() - 1; }
        if (b.ring >= 0) { rgs.push_back(from.rings()[(size_t)b.ring]); b.ring = (int64_t)rgs.size() - 1; }
        m.push_back(b);
    }
    void clear() { m.clear(); gps.clear(); rgs.clear(); }
};

int64_t mid_in(const MSystem &s, uint64_t sid) { return s.get_mid(sid); }

}  // namespace

int MSystem::combined_integrate(double dt, int64_t n_combine, int64_t n_step, int use_gpu) {
    if (mlist.empty()) { err = "no system loaded"; return fail(EPH_EINVAL, err); }
    if (!(dt == dt) || dt == 0 || n_combine <= 0) { err = "dt must be finite and non-zero, n_combine > 0"; return fail(EPH_EINVAL, err); }
    if (use_gpu != 0 && use_gpu != 1) { err = "USE_GPU must be 0 or 1"; return fail(EPH_EINVAL, err); }
    analyse();
    if (!analysis_settled || blist.empty()) {
        err = "combined integration needs the orbit structure, and its analysis did not settle";
        return fail(EPH_EINVAL, err);
    }
    const int64_t bn = (int64_t)blist.size();
    std::map<int64_t, int64_t> clist;                    // child -> parent
    for (int64_t id = 0; id < bn; ++id) {
        const Barycen &b = blist[(size_t)id];
        if (b.hid >= 0) continue;
        const Barycen &t = blist[(size_t)b.tid];
        if (t.pid < 0) continue;
        const Barycen &tp = blist[(size_t)t.pid];
        if (todbl(t.GM_sys) > GM_max_child || todbl(tp.GM) > GM_max_parent) continue;
        clist.insert({b.mid, tp.mid});
    }
    for (;;) {                                           // chains resolve to the top parent
        bool changed = false;
        for (auto &c : clist) {
            auto fs = clist.find(c.second);
            if (fs == clist.end()) continue;
            c.second = fs->second;
            changed = true;
        }
        if (!changed) break;
    }
    std::map<int64_t, std::vector<int64_t>> cvecs;       // parent -> children
    for (const auto &c : clist) cvecs[c.second].push_back(c.first);

    auto minor = [&](double GM) { return GM < GM_max_tiny; };     // IS_MINOR
    std::map<int64_t, std::unique_ptr<MSystem>> Sn;
    for (const auto &p : cvecs) Sn[p.first].reset(new MSystem());
    MSystem Sc, Sx;
    for (MSystem *s : {&Sc, &Sx}) s->n_threads = n_threads;
    for (auto &sn : Sn) sn.second->n_threads = 1;      // subsystems run in parallel, one thread each
    std::vector<int64_t> Smajsub;
    const int64_t mn = size();
    const double dt_long = dt * (double)n_combine;
    SubBuilder bc, bx;
    std::map<int64_t, SubBuilder> bn_;
    std::map<int64_t, std::vector<int64_t>> sn_tidal_children;
    std::map<int64_t, int64_t> sn_tidal_parent;

    for (int64_t i_step = 0; i_step < n_step; ++i_step) {
        bc.clear();
        bx.clear();
        bn_.clear();
        sn_tidal_children.clear();
        sn_tidal_parent.clear();
        for (const auto &sn : Sn) bn_[sn.first];
        Smajsub.clear();
        for (int64_t i = 0; i < mn; ++i) {
            const eph_mass &mi = mlist[(size_t)i];
            auto clp = clist.find(i);
            if (clp != clist.end()) {                    // a child: only in its parent's subsystem
                SubBuilder &sb = bn_[clp->second];
                sn_tidal_children[clp->second].push_back((int64_t)sb.m.size());
                sb.push(mi, *this);
                continue;
            }
            auto cvp = cvecs.find(i);
            if (cvp != cvecs.end()) {                    // a parent: combined with its children
                const auto &chs = cvp->second;
                eph_mass mci = mi;
                mci.gp = -1;
                if (mci.ring >= 0) {
                    const eph_ring &mr = rgs[(size_t)mci.ring];
                    mci.ring = -1;
                    double rGM = mr.GM_ratio;
                    rGM /= 1 - rGM;
                    mci.GM += mci.GM * rGM;
                    mci.GM0 += mci.GM * rGM;
                    const v3 mgl = todbl(DV(mci.GL));
                    const v3 gx = asc_node(mgl), gz = unit(mgl);
                    const m3 fgls{gx, cross(gz, gx), gz};
                    m3 rh = mzero();
                    rh.x.x = mr.J2 / 2;
                    rh.y.y = mr.J2 / 2;
                    rh.z.z = -mr.J2;
                    m3 cp = M(mci.C_potential) + toworld(fgls, rh);
                    cp = cp - mr.GM_ratio * cp;
                    mci.C_potential = E(cp);
                }
                mci.k2 = 0;                              // parent rotation frozen, angular momentum kept
                mci.k2r = 0;
                mci.dJ2 = 0;
                mci.C_static = mci.C_potential;
                mci.s = E(mkddm(m3{v3{1, 0, 0}, v3{0, 1, 0}, v3{0, 0, 1}}));
                mci.A *= kVeryBig;
                mci.inertia *= kVeryBig;
                mci.w = E(V(mci.w) / kVeryBig);
                dd GM = mkdd(mci.GM), GM0 = mkdd(mci.GM0);
                ddv3 r = eph::scale(GM, DV(mi.r)), v = eph::scale(GM, DV(mi.v));
                for (int64_t j : chs) {
                    const eph_mass &mj = mlist[(size_t)j];
                    const dd GMj = mkdd(mj.GM);
                    r = add(r, eph::scale(GMj, DV(mj.r)));
                    v = add(v, eph::scale(GMj, DV(mj.v)));
                    GM = add(GM, GMj);
                    GM0 = add(GM0, mkdd(mj.GM0));
                    mci.dGM += mj.dGM;
                    mci.lum += mj.lum;
                    // CombineCorrector: the child's orbital angular momentum and
                    // the averaged quadrupole of its orbit over the long step
                    const v3 dr = todbl(sub(DV(mj.r), DV(mci.r)));
                    v3 ox = dr;                          // mat_t(dr, orthogonalizer(dv), 0), vec_t.h:284-300
                    const v3 dvv = todbl(sub(DV(mj.v), DV(mci.v)));
                    v3 oz = cross(ox, dvv);
                    const double axis_norm = normalize(ox);
                    const v3 xx = asc_node(ox), xy = cross(ox, xx);
                    double zx = dot(xx, oz), zy = dot(xy, oz);
                    const double cross_norm = normalize2(zy, zx);
                    oz = zx * xx + zy * xy;
                    const v3 oy = zy * xx - zx * xy;
                    const m3 mrot{ox, oy, oz};
                    const v3 dj = oz * cross_norm;
                    mci.GL = E(add(DV(mci.GL), mkddv((todbl(DV(mj.GL)) + dj) * (mj.GM / mi.GM))));
                    const double r2 = axis_norm * axis_norm;
                    const double dw = cross_norm / r2;
                    const double t = dt_long * dw;
                    const double st = std::sin(t), ct = std::cos(t), sit = t == 0 ? 3 : 3 * (st / t);
                    m3 dh{v3{(-1 - sit * ct) / 2, -sit * st / 2, 0}, v3{-sit * st / 2, (-1 + sit * ct) / 2, 0},
                          v3{0, 0, 1}};
                    dh = toworld(mrot, dh);
                    double dhnorm = mj.GM + mci.GM;
                    dhnorm = -r2 * mj.GM * mci.GM / (dhnorm * dhnorm * 2 * mci.R2);
                    dh = dhnorm * dh;
                    mci.C_static = E(M(mci.C_static) + dh);
                }
                const dd rGMs = div(mkdd(1), GM);
                mci.r = E(eph::scale(rGMs, r));
                mci.v = E(eph::scale(rGMs, v));
                mci.GM0 = todbl(GM0);
                mci.GM = todbl(GM);
                bc.push(mci, *this);
                SubBuilder &own = bn_[i];                // the parent itself in its own subsystem
                sn_tidal_parent[i] = (int64_t)own.m.size();
                own.push(mi, *this);
                if (!minor(mci.GM0)) {
                    bx.push(mci, *this);
                    Smajsub.push_back(i);
                    for (auto &sb : bn_)                 // the combined parent in every other subsystem
                        if (sb.first != i) sb.second.push(mci, *this);
                } else {
                    scale_mass(mci, 1 / kVeryBig);       // a tiny parent is only a test mass in Sx
                    bx.push(mci, *this);
                }
                continue;
            }
            bc.push(mi, *this);                          // neither parent nor child
            if (!minor(mi.GM0)) {
                bx.push(mi, *this);
                for (auto &sb : bn_) sb.second.push(mi, *this);
            }
        }
        int rc;
        if ((rc = Sc.assign((int64_t)bc.m.size(), bc.m.data(), bc.gps.data(), (int64_t)bc.gps.size(), bc.rgs.data(),
                            (int64_t)bc.rgs.size(), t_eph_)))
            return rc;
        if ((rc = Sc.accel(0))) return rc;
        if ((rc = Sx.assign((int64_t)bx.m.size(), bx.m.data(), bx.gps.data(), (int64_t)bx.gps.size(), bx.rgs.data(),
                            (int64_t)bx.rgs.size(), t_eph_)))
            return rc;
        if ((rc = Sx.accel(0))) return rc;
        for (auto &sns : Sn) {
            MSystem &sn = *sns.second;
            SubBuilder &sb = bn_[sns.first];
            if ((rc = sn.assign((int64_t)sb.m.size(), sb.m.data(), sb.gps.data(), (int64_t)sb.gps.size(),
                                sb.rgs.data(), (int64_t)sb.rgs.size(), t_eph_)))
                return rc;
            m3 tm = mzero();                             // tidal field of the tiny bodies of Sc
            const eph_mass &mp = Sc.mlist[(size_t)mid_in(Sc, mlist[(size_t)sns.first].sid)];
            for (const eph_mass &mj : Sc.mlist) {
                if (!(minor(mj.GM0) && mj.sid != mp.sid)) continue;
                v3 rm = todbl(sub(DV(mj.r), DV(mp.r)));
                const v3 vm = todbl(sub(DV(mj.v), DV(mp.v)));
                rm += vm * (dt_long / 2);
                const double rr2 = 1 / smax(dot(rm, rm), mj.sR2), rr = sqrt(rr2);
                const double tmatn = -mj.GM0 * rr2 * rr;
                tm = tm - subdiag(outer((3 * tmatn * rr2) * rm, rm), tmatn);
            }
            if ((rc = sn.set_tidal(sn_tidal_parent[sns.first], tm, sn_tidal_children[sns.first]))) return rc;
            if ((rc = sn.accel(0))) return rc;
        }
        // integrate: the subsystems' short steps on worker threads, then the
        // test-mass system and the full combined system
        std::vector<int> sub_rc(Sn.size(), 0);
        {
            std::vector<std::thread> th;
            int q = 0;
            for (auto &sns : Sn) {
                MSystem *sn = sns.second.get();
                int *out = &sub_rc[(size_t)q++];
                th.emplace_back([sn, dt, n_combine, out] { *out = sn->integrate(dt, n_combine, 0); });
            }
            rc = Sx.integrate(dt_long, 1, 0);
            const int rc2 = rc ? rc : Sc.integrate(dt_long, 1, use_gpu);
            for (auto &t : th) t.join();
            if (rc2) return rc2;
            for (int r : sub_rc)
                if (r) return r;
        }
        // finalize: merge the subsystems' motion into the full system
        for (int64_t i = 0; i < mn; ++i) {
            eph_mass &mi = mlist[(size_t)i];
            if (clist.count(i)) continue;
            auto cvp = cvecs.find(i);
            if (cvp != cvecs.end()) {
                const auto &chs = cvp->second;
                const eph_mass &Sci = Sc.mlist[(size_t)mid_in(Sc, mi.sid)];
                const eph_mass &Sxi = Sx.mlist[(size_t)mid_in(Sx, mi.sid)];
                ddv3 dr = sub(DV(Sci.r), DV(Sxi.r)), dv = sub(DV(Sci.v), DV(Sxi.v));
                if (!minor(Sci.GM0)) {
                    for (int64_t si : Smajsub) {
                        if (si == i) continue;
                        const MSystem &Sj = *Sn[si];
                        const eph_mass &Sji = Sj.mlist[(size_t)mid_in(Sj, mi.sid)];
                        dr = add(dr, sub(DV(Sji.r), DV(Sxi.r)));
                        dv = add(dv, sub(DV(Sji.v), DV(Sxi.v)));
                    }
                }
                const MSystem &Si = *Sn[i];
                mass_copy(mi, Si.mlist[(size_t)mid_in(Si, mi.sid)]);
                mi.r = E(add(DV(mi.r), dr));
                mi.v = E(add(DV(mi.v), dv));
                for (int64_t j : chs) {
                    eph_mass &mj = mlist[(size_t)j];
                    mass_copy(mj, Si.mlist[(size_t)mid_in(Si, mj.sid)]);
                    mj.r = E(add(DV(mj.r), dr));
                    mj.v = E(add(DV(mj.v), dv));
                }
                continue;
            }
            const eph_mass &Sci = Sc.mlist[(size_t)mid_in(Sc, mi.sid)];
            mass_copy(mi, Sci);
            if (!minor(mi.GM0)) {
                ddv3 dr = kZero3, dv = kZero3;
                const eph_mass &Sxi = Sx.mlist[(size_t)mid_in(Sx, mi.sid)];
                for (int64_t si : Smajsub) {
                    const MSystem &Sj = *Sn[si];
                    const eph_mass &Sji = Sj.mlist[(size_t)mid_in(Sj, mi.sid)];
                    dr = add(dr, sub(DV(Sji.r), DV(Sxi.r)));
                    dv = add(dv, sub(DV(Sji.v), DV(Sxi.v)));
                }
                mi.r = E(add(DV(mi.r), dr));
                mi.v = E(add(DV(mi.v), dv));
            }
            // rotations too fast for the long step are advanced about GL
            if (length(V(mi.w)) * std::abs(dt_long) > 2 * kPi / kRotationStepLimit) {
                const ddm3 s = DM(mi.s);
                v3 zl = todbl(cross(s.z, DV(mi.GL)));
                const double l = length(todbl(DV(mi.GL)));
                if (l > 0) {
                    zl = zl / l;
                    mi.s = E(add(s, rotation_matrix(zl, 1.0, [](double x) { return std::sin(x); }) * todbl(s)));
                }
            }
        }
        t_eph_ = Sc.t_eph_;
        update(todbl(t_eph_));
        int rc3 = accel(use_gpu);
        if (rc3) return rc3;
    }
    host_changed();
    return EPH_OK;
}

// ---------------------------------------------------------------------------
// checkpoints: msystem::save_checkpoint / load_checkpoint
// (src/modules/mass_loader.cpp:566-759), magic and version of
// src/configs.cpp:5-6
// ---------------------------------------------------------------------------
namespace {

constexpr int64_t kCheckPointMagic = 0x53484c486d687045;
constexpr int64_t kCheckPointVersion = 4;

// fcp_header (mass_loader.cpp:569-580): 17 eight-byte fields
struct CheckpointHeader {
    int64_t magic, version;
    double t_eph_hi, t_eph_lo;
    double delta_t;
    int64_t integrator;
    double data_cadence, max_ephm_length, combined_delta_t;
    double GM_max_child, GM_max_parent, GM_max_tiny, Period_max_child;
    int64_t nmass, nbarycen, masssize;
};
static_assert(sizeof(CheckpointHeader) == 16 * 8, "fcp_header is sixteen 8-byte fields");
struct BarycenIds { int64_t pid, hid, gid, tid, mid, nch; };   // barycen_ids, mass_loader.cpp:581-584

// geopotential object bytes: {Nz, Nt, c_table[Nz (Nz + 2) - 3]} (geopotential.cpp:133-135)
size_t gp_bytes(const eph_geopotential &g) { return 16 + 8 * (size_t)(g.Nz * (g.Nz + 2) - 3); }
// ring object bytes: {N, GM_ratio, A, J2, GL_R2, {GsR2, R_R, H_R}[N]} (ring.cpp:114-116)
size_t ring_bytes(const eph_ring &r) { return 40 + 24 * (size_t)r.N; }

struct Reader {
    const char *p, *end;
    bool get(void *dst, size_t n) {
        if ((size_t)(end - p) < n) return false;
        memcpy(dst, p, n);
        p += n;
        return true;
    }
};

// bsystem::compatible_size (mass_loader.cpp:828-877): the number of bodies
// of a well-formed hierarchy (one root, every body once, binaries with host
// and guest, tid and mid consistent), 0 when it is not one
int64_t compatible_size(const std::vector<Barycen> &bl) {
    const int64_t bn = (int64_t)bl.size();
    int64_t root = -1, nroot = 0;
    for (int64_t i = 0; i < bn; ++i)
        if (bl[(size_t)i].pid < 0) { root = i; ++nroot; }
    if (nroot != 1 || root < 0) return 0;
    int64_t mn = bn;
    for (const Barycen &b : bl) mn -= b.hid >= 0;
    std::vector<char> has_done((size_t)bn, 0), has_mass((size_t)mn, 0), mid_done((size_t)bn, 0);
    std::vector<std::pair<int64_t, int64_t>> active{{root, -1}};
    for (size_t head = 0; head < active.size(); ++head) {
        const int64_t bid = active[head].first, from = active[head].second;
        if (bid < 0 || bid >= bn || has_done[(size_t)bid]) return 0;
        has_done[(size_t)bid] = 1;
        const Barycen &b = bl[(size_t)bid];
        if (from != b.pid) return 0;
        const bool is_barycen = b.hid >= 0;
        if (is_barycen != (b.gid >= 0)) return 0;
        if (is_barycen) {
            active.emplace_back(b.hid, bid);
            active.emplace_back(b.gid, bid);
        } else {
            if (b.mid < 0 || b.mid >= mn || has_mass[(size_t)b.mid] || mid_done[(size_t)bid]) return 0;
            has_mass[(size_t)b.mid] = 1;
            mid_done[(size_t)bid] = 1;
        }
        int64_t ctid = bid;
        if (b.pid >= 0) {
            const Barycen &p = bl[(size_t)b.pid];
            if (p.hid == bid) {
                ctid = p.tid;
                if (b.mid != p.mid || mid_done[(size_t)b.pid]) return 0;
                mid_done[(size_t)b.pid] = 1;
            }
        }
        if (b.tid != ctid) return 0;
        for (int64_t c : b.children) active.emplace_back(c, bid);
    }
    for (char c : has_done) if (!c) return 0;
    for (char c : mid_done) if (!c) return 0;
    for (char c : has_mass) if (!c) return 0;
    return mn;
}

}  // namespace

int MSystem::save_checkpoint(std::vector<char> &out) const {
    out.clear();
    auto put = [&](const void *src, size_t n) { out.insert(out.end(), (const char *)src, (const char *)src + n); };
    CheckpointHeader h;
    h.magic = kCheckPointMagic;
    h.version = kCheckPointVersion;
    h.t_eph_hi = t_eph_.hi;
    h.t_eph_lo = t_eph_.lo;
    h.delta_t = delta_t_;
    h.integrator = integrator_;
    h.data_cadence = data_cadence;
    h.max_ephm_length = max_ephm_length;
    h.combined_delta_t = combined_delta_t;
    h.GM_max_child = GM_max_child;
    h.GM_max_parent = GM_max_parent;
    h.GM_max_tiny = GM_max_tiny;
    h.Period_max_child = Period_max_child;
    h.nmass = size();
    h.nbarycen = (int64_t)blist.size();
    h.masssize = (int64_t)kCheckpointMassBytes;
    put(&h, sizeof h);
    for (const Barycen &b : blist) {                     // bsystem::save_barycen_structure, mass_loader.cpp:700-713
        const BarycenIds ids{b.pid, b.hid, b.gid, b.tid, b.mid, (int64_t)b.children.size()};
        put(&ids, sizeof ids);
        put(b.children.data(), sizeof(int64_t) * b.children.size());
    }
    for (const eph_mass &m : mlist) {                    // model pointers become the models' byte counts
        eph_mass w = m;
        w.gp = m.gp >= 0 ? (int64_t)gp_bytes(gps[(size_t)m.gp]) : 0;
        w.ring = m.ring >= 0 ? (int64_t)ring_bytes(rgs[(size_t)m.ring]) : 0;
        put(&w, kCheckpointMassBytes);
        if (m.gp >= 0) {
            const eph_geopotential &g = gps[(size_t)m.gp];
            put(&g.Nz, 8);
            put(&g.Nt, 8);
            put(g.c, gp_bytes(g) - 16);
        }
        if (m.ring >= 0) {
            const eph_ring &r = rgs[(size_t)m.ring];
            put(&r.N, 8);
            put(&r.GM_ratio, 8);
            put(&r.A, 8);
            put(&r.J2, 8);
            put(&r.GL_R2, 8);
            for (int64_t q = 0; q < r.N; ++q) {
                put(&r.GsR2[q], 8);
                put(&r.R_R[q], 8);
                put(&r.H_R[q], 8);
            }
        }
    }
    return EPH_OK;
}

int MSystem::save_checkpoint(const std::string &path) const {
    std::vector<char> buf;
    int rc = save_checkpoint(buf);
    if (rc) return rc;
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return fail(EPH_EIO, path + ": cannot open for writing");
    const bool ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (fclose(f) != 0 || !ok) return fail(EPH_EIO, path + ": write failed");
    return EPH_OK;
}

int MSystem::load_checkpoint(const char *data, size_t size_b) {
    Reader rd{data, data + size_b};
    CheckpointHeader h;
    auto bad = [&](const std::string &why) { err = "checkpoint: " + why; return fail(EPH_EIO, err); };
    if (!rd.get(&h, sizeof h) || h.magic != kCheckPointMagic || h.nmass <= 0 || h.nbarycen <= 0)
        return bad("not a checkpoint");
    if (h.version != kCheckPointVersion || h.masssize != (int64_t)kCheckpointMassBytes)
        return bad("version of the checkpoint differs from the program's");
    std::vector<Barycen> bl;
    for (int64_t i = 0; i < h.nbarycen; ++i) {           // bsystem::load_barycen_structure, mass_loader.cpp:586-612
        BarycenIds ids;
        if (!rd.get(&ids, sizeof ids) || ids.nch < 0 || ids.nch > h.nbarycen) return bad("truncated barycen list");
        Barycen b;
        b.pid = ids.pid; b.hid = ids.hid; b.gid = ids.gid; b.tid = ids.tid; b.mid = ids.mid;
        b.children.resize((size_t)ids.nch);
        if (ids.nch && !rd.get(b.children.data(), sizeof(int64_t) * (size_t)ids.nch)) return bad("truncated barycen list");
        bl.push_back(std::move(b));
    }
    std::vector<eph_mass> ms;
    std::vector<eph_geopotential> g;
    std::vector<eph_ring> r;
    for (int64_t i = 0; i < h.nmass; ++i) {
        eph_mass m;
        memset(&m, 0, sizeof m);
        if (!rd.get(&m, kCheckpointMassBytes)) return bad("truncated mass list");
        if (m.gp) {
            eph_geopotential gp;
            memset(&gp, 0, sizeof gp);
            if (!rd.get(&gp.Nz, 8) || !rd.get(&gp.Nt, 8)) return bad("truncated geopotential");
            if (gp.Nz < 2 || gp.Nz > EPH_GP_MAX_DEGREE || gp.Nt < 0 || gp.Nt > gp.Nz || (int64_t)gp_bytes(gp) != m.gp)
                return bad("geopotential size does not match its degree");
            if (!rd.get(gp.c, gp_bytes(gp) - 16)) return bad("truncated geopotential");
            m.gp = (int64_t)g.size();
            g.push_back(gp);
        } else
            m.gp = -1;
        if (m.ring) {
            eph_ring rg;
            memset(&rg, 0, sizeof rg);
            if (!rd.get(&rg.N, 8) || rg.N < 0 || rg.N > EPH_RING_MAX_DISKS || (int64_t)ring_bytes(rg) != m.ring)
                return bad("ring size does not match its disk count");
            if (!rd.get(&rg.GM_ratio, 8) || !rd.get(&rg.A, 8) || !rd.get(&rg.J2, 8) || !rd.get(&rg.GL_R2, 8))
                return bad("truncated ring");
            for (int64_t q = 0; q < rg.N; ++q)
                if (!rd.get(&rg.GsR2[q], 8) || !rd.get(&rg.R_R[q], 8) || !rd.get(&rg.H_R[q], 8)) return bad("truncated ring");
            m.ring = (int64_t)r.size();
            r.push_back(rg);
        } else
            m.ring = -1;
        ms.push_back(m);
    }
    int rc = assign((int64_t)ms.size(), ms.data(), g.data(), (int64_t)g.size(), r.data(), (int64_t)r.size(),
                    dd{h.t_eph_hi, h.t_eph_lo});
    if (rc) return rc;
    delta_t_ = h.delta_t;
    integrator_ = h.integrator;
    data_cadence = h.data_cadence;
    max_ephm_length = h.max_ephm_length;
    combined_delta_t = h.combined_delta_t;
    GM_max_child = h.GM_max_child;
    GM_max_parent = h.GM_max_parent;
    GM_max_tiny = h.GM_max_tiny;
    Period_max_child = h.Period_max_child;
    blist.swap(bl);
    // a structure that fits these bodies gets its state back (mass_loader.cpp:685-688)
    if (compatible_size(blist) == h.nmass) {
        const int64_t n_b = (int64_t)blist.size();
        for (auto &b : blist) { b.GM_sys = dd{0, 0}; b.r_sys = b.v_sys = kZero3; }
        for (int64_t i = 0; i < n_b; ++i) {              // bsystem::update_barycens, mass_loader.cpp:335-372
            Barycen &bi = blist[(size_t)i];
            if (bi.hid >= 0) continue;
            const eph_mass &mi = mlist[(size_t)bi.mid];
            const dd mGM = mkdd(mi.GM);
            const ddv3 rGM = eph::scale(mGM, DV(mi.r)), vGM = eph::scale(mGM, DV(mi.v));
            bi.r = DV(mi.r);
            bi.v = DV(mi.v);
            bi.GM = mGM;
            bi.r_sys = add(bi.r_sys, rGM);
            bi.v_sys = add(bi.v_sys, vGM);
            bi.GM_sys = add(bi.GM_sys, mGM);
            for (int64_t bp = bi.pid; bp >= 0; bp = blist[(size_t)bp].pid) {
                Barycen &p = blist[(size_t)bp];
                p.GM_sys = add(p.GM_sys, mGM);
                p.r_sys = add(p.r_sys, rGM);
                p.v_sys = add(p.v_sys, vGM);
            }
        }
        for (auto &bi : blist) {
            if (bi.hid < 0) continue;
            const Barycen &hh = blist[(size_t)bi.hid], &gg = blist[(size_t)bi.gid];
            bi.GM = add(hh.GM_sys, gg.GM_sys);
            bi.r = divide(add(hh.r_sys, gg.r_sys), bi.GM);
            bi.v = divide(add(hh.v_sys, gg.v_sys), bi.GM);
        }
        for (auto &bi : blist) {
            bi.r_sys = divide(bi.r_sys, bi.GM_sys);
            bi.v_sys = divide(bi.v_sys, bi.GM_sys);
        }
        t_update = t_eph_;
    }
    return EPH_OK;
}

int MSystem::load_checkpoint(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) { err = path + ": cannot open"; return fail(EPH_EIO, err); }
    std::vector<char> buf;
    char tmp[1 << 16];
    size_t got;
    while ((got = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
    fclose(f);
    return load_checkpoint(buf.data(), buf.size());
}

}  // namespace eph
