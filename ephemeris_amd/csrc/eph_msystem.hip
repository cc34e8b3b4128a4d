// Text-format loader and the msystem-style host class (see eph_msystem.hpp).
#include "eph_msystem.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include <algorithm>
#include <thread>

#include "eph_cpu.hpp"
#include "eph_math.hpp"
#include "eph_runtime.hpp"

namespace eph {

namespace {

constexpr double kEps = 0x1p-53;                     // basic_math.h:28
constexpr double kDegree = kPi / 180;
constexpr double kObliquity = 84381.448 / 3600 * kDegree;

// MFILE::readline (src/utils/memio.cpp:257-301): '\r' and '\n' both end a
// line; empty lines and lines whose first character is '#' are skipped.
struct LineReader {
    FILE *f = nullptr;
    explicit LineReader(const std::string &path) { f = fopen(path.c_str(), "rb"); }
    ~LineReader() { if (f) fclose(f); }
    bool next(std::string &line) {
        for (;;) {
            line.clear();
            int c;
            bool got = false;
            while ((c = fgetc(f)) != EOF) {
                got = true;
                if (c == '\r' || c == '\n') break;
                line.push_back((char)c);
            }
            if (!line.empty() && line[0] == '#') line.clear();
            if (!line.empty()) return true;
            if (!got || c == EOF) return false;
        }
    }
};

v3 V(const eph_vec &a) { return v3{a.x, a.y, a.z}; }
eph_vec E(v3 a) { return eph_vec{a.x, a.y, a.z}; }
m3 M(const eph_mat &a) { return m3{V(a.x), V(a.y), V(a.z)}; }
eph_mat E(const m3 &a) { return eph_mat{E(a.x), E(a.y), E(a.z)}; }
ddv3 DV(const eph_ddvec &a) { return ddv3{dd{a.x.hi, a.x.lo}, dd{a.y.hi, a.y.lo}, dd{a.z.hi, a.z.lo}}; }
eph_ddvec E(const ddv3 &a) { return eph_ddvec{{a.x.hi, a.x.lo}, {a.y.hi, a.y.lo}, {a.z.hi, a.z.lo}}; }
ddm3 DM(const eph_ddmat &a) { return ddm3{DV(a.x), DV(a.y), DV(a.z)}; }
eph_ddmat E(const ddm3 &a) { return eph_ddmat{E(a.x), E(a.y), E(a.z)}; }

struct HostSin { double operator()(double x) const { return std::sin(x); } };

// geopotential::load (src/physics/geopotential.cpp:8-127), N_start = 3
int load_gp(const std::string &path, double factor, eph_geopotential &out, std::string &err) {
    LineReader rd(path);
    if (!rd.f) { err = path + ": cannot open geopotential model"; return EPH_EIO; }
    std::string line;
    long long Nz = -1, Nt = 0;
    int64_t Nzmax = 0, Ntmax = 0;
    const int64_t N_start = 3;
    bool have_header = false;
    memset(&out, 0, sizeof out);
    while (rd.next(line)) {
        if (!have_header) {
            if (sscanf(line.c_str(), "%lld%lld", &Nz, &Nt) != 2) { err = path + ": bad degree header"; return EPH_EIO; }
            if (Nt < 0) Nt = 0;
            if (Nz < Nt) Nz = Nt;
            if (Nz > EPH_GP_MAX_DEGREE) { Nz = EPH_GP_MAX_DEGREE; if (Nt > EPH_GP_MAX_DEGREE) Nt = EPH_GP_MAX_DEGREE; }
            if (Nz < N_start) { err = path + ": geopotential degree too low"; return EPH_EIO; }
            out.Nz = Nz;
            out.Nt = Nt;
            have_header = true;
            continue;
        }
        long long n = 0, m = 0;
        double Cnm = 0, Snm = 0;
        int nr = sscanf(line.c_str(), "%lld%lld%lf%lf", &n, &m, &Cnm, &Snm);
        if (nr < 3 || n < 2 || m < 0 || m > n || (nr == 3 && m != 0) || (m == 0 && Snm != 0)) {
            err = path + ": malformed term: " + line;
            return EPH_EIO;
        }
        if (n < N_start) continue;
        if (n > Nz || (n > Nt && m > 0)) continue;
        double rr = std::pow(factor, (double)n);
        Cnm *= rr;
        Snm *= rr;
        const int64_t J = n * n - 4;
        if (m == 0) out.c[J] += Cnm;
        else {
            out.c[J + 2 * m - 1] += Cnm;
            out.c[J + 2 * m] += Snm;
            Ntmax = std::max<int64_t>(Ntmax, n);
        }
        Nzmax = std::max<int64_t>(Nzmax, n);
    }
    if (!have_header) { err = path + ": empty geopotential model"; return EPH_EIO; }
    if (!Nzmax) { err = path + ": no valid geopotential terms"; return EPH_EIO; }
    if (Nzmax < out.Nz) out.Nz = Nzmax;
    if (Ntmax < out.Nt) out.Nt = Ntmax;
    return EPH_OK;
}

// ring::load (src/physics/ring.cpp:25-108)
int load_ring(const std::string &path, double ref_GM, double ref_R, double dmf, eph_ring &out, std::string &err) {
    LineReader rd(path);
    if (!rd.f) { err = path + ": cannot open ring model"; return EPH_EIO; }
    std::string line;
    std::vector<double> Gs, Rs, Hs;
    std::vector<char> name(1100);
    while (rd.next(line)) {
        if (line.size() >= 1024) { err = path + ": line too long"; return EPH_EIO; }
        double rin, rout, thick, GM;
        if (sscanf(line.c_str(), "%[^\t]%lf%lf%lf%lf", name.data(), &rin, &rout, &thick, &GM) != 5) {
            err = path + ": malformed ring line: " + line;
            return EPH_EIO;
        }
        GM *= std::abs(dmf);
        rin *= 1000; rout *= 1000; thick *= 1000;
        const double g = GM / (kPi * (rout - rin) * (rout + rin));
        const double add_g[2] = {-g, g}, add_r[2] = {rin, rout};
        for (int q = 0; q < 2; ++q) {                 // adddisk, ring.cpp:15-23
            bool found = false;
            for (size_t d = 0; d < Rs.size(); ++d)
                if (Rs[d] == add_r[q] && Hs[d] == thick) { Gs[d] += add_g[q]; found = true; break; }
            if (!found) { Gs.push_back(add_g[q]); Rs.push_back(add_r[q]); Hs.push_back(thick); }
        }
    }
    if ((int64_t)Rs.size() > EPH_RING_MAX_DISKS) { err = path + ": too many disks"; return EPH_EINVAL; }
    memset(&out, 0, sizeof out);
    out.N = (int64_t)Rs.size();
    const double ref_R2 = ref_R * ref_R, ref_GMR2 = ref_GM * ref_R2;
    for (int64_t i = 0; i < out.N; ++i) {
        const double R2 = Rs[i] * Rs[i], H2 = Hs[i] * Hs[i];
        const double GM = kPi * R2 * Gs[i];
        out.GM_ratio += GM;
        out.A += GM * (6 * R2 + H2) / (12 * ref_GMR2);
        out.J2 += GM * (3 * R2 - H2) / (12 * ref_GMR2);
        out.GL_R2 += 4 * GM * std::sqrt(ref_GM * Rs[i]) / (5 * ref_GMR2);
        out.GsR2[i] = Gs[i] / ref_GM * ref_R2;
        out.R_R[i] = Rs[i] / ref_R;
        out.H_R[i] = Hs[i] / ref_R;
    }
    out.GM_ratio = 1 / (1 + ref_GM / out.GM_ratio);
    if (dmf < 0) out.GL_R2 *= -1;
    return EPH_OK;
}

std::string dir_of(const std::string &p) {
    const size_t s = p.find_last_of("/\\");
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

}  // namespace

bool mass_sanity(const eph_mass &m, std::string *why) {   // mass_loader.cpp:18-65
    const double eps2 = kEps * kEps, large_finite = 1 / eps2;
    std::vector<const char *> errs;
    const v3 r = todbl(DV(m.r)), v = todbl(DV(m.v));
    if (!(dot(r, r) < large_finite * large_finite) || !(dot(v, v) < kC2)) errs.push_back("position/velocity not finite");
    const m3 f = todbl(DM(m.s));
    const m3 o = adddiag(-(transpose(f) * f), 1);
    if (!(dot(o.x, o.x) + dot(o.y, o.y) + dot(o.z, o.z) < 64 * eps2) || !(det(f) > 0)) errs.push_back("rotation not normalized");
    if (!(dot(V(m.w), V(m.w)) * m.R2 < kC2)) errs.push_back("rotation too fast");
    if (!(eps2 < m.GM0 && m.GM0 < large_finite)) errs.push_back("invalid gravitational parameter");
    if (!(2 * m.GM0 < kC2 * std::abs(m.R))) errs.push_back("invalid radius");
    if (!(eps2 < m.inertia && m.inertia <= 1)) errs.push_back("invalid inertia factor");
    if (!(0 <= m.k2 && m.k2 <= 1.5 && 0 <= m.k2r && m.k2r <= 1.5)) errs.push_back("invalid tidal/rotational deformation factor");
    if (!(0 <= m.tide_delay && m.tide_delay < large_finite && 0 <= m.tide_delay_factor && m.tide_delay_factor <= large_finite))
        errs.push_back("invalid tidal delay time/factor");
    const m3 C = M(m.C_static);                          // mat_t::to_harmonics, vec_t.h:531-537
    const double j2 = -C.z.z, c21 = (C.x.z + C.z.x) / 3, c22 = (C.x.x - C.y.y) / 6;
    const double s21 = (C.y.z + C.z.y) / 3, s22 = (C.x.y + C.y.x) / 6;
    if (!(0 <= j2 && j2 + 6 * std::abs(c22) < 3 * m.inertia && std::abs(c21) < m.inertia && std::abs(s21) < m.inertia &&
          std::abs(s22) * 2 < m.inertia))
        errs.push_back("invalid static harmonics");
    if (m.R < 0 && (m.k2 || m.k2r || j2 || m.dJ2 || c21 || c22 || s21 || s22 || m.gp >= 0 || m.ring >= 0 || m.tide_delay))
        errs.push_back("soft mass shall not have non-point components");
    if (why)
        for (const char *e : errs) { *why += e; *why += "; "; }
    return errs.empty();
}

MSystem::MSystem() = default;

MSystem::~MSystem() {
    drop_device();
}

void MSystem::drop_device() {
    if (sys) eph_system_destroy(sys);
    sys = nullptr;
    gpu_stale = true;
}

void MSystem::reset_params() {                            // msystem::reset_params, mass_loader.cpp:296-308
    t_eph_ = dd{0, 0};
    delta_t_ = 300;
    integrator_ = COMBINED_RK12;
    data_cadence = 7200;
    max_ephm_length = 631152000;
    combined_delta_t = 28800;
    GM_max_parent = 2E17;
    GM_max_child = 1E13;
    GM_max_tiny = 14E11;
    Period_max_child = 0;
}

void MSystem::copy_params(const MSystem &o) {            // msystem::copy_params, mass_loader.cpp:924-937
    t_eph_ = o.t_eph_;
    delta_t_ = o.delta_t_;
    integrator_ = o.integrator_;
    data_cadence = o.data_cadence;
    max_ephm_length = o.max_ephm_length;
    combined_delta_t = o.combined_delta_t;
    GM_max_child = o.GM_max_child;
    GM_max_parent = o.GM_max_parent;
    GM_max_tiny = o.GM_max_tiny;
    Period_max_child = o.Period_max_child;
}

void MSystem::params(double *o) const {
    o[0] = delta_t_; o[1] = data_cadence; o[2] = max_ephm_length; o[3] = combined_delta_t;
    o[4] = GM_max_child; o[5] = GM_max_parent; o[6] = GM_max_tiny; o[7] = Period_max_child;
}

WorkerPool *MSystem::pool() {
    if (!pool_) {
        int n = n_threads;
        if (n <= 0) n = (int)std::max(1u, std::thread::hardware_concurrency());
        pool_.reset(new WorkerPool(n));
    }
    return pool_.get();
}

int MSystem::set_threads(int n) {
    if (n < 0) { err = "thread count must be >= 0"; return fail(EPH_EINVAL, err); }
    n_threads = n;
    pool_.reset();
    return EPH_OK;
}

int MSystem::threads() const { return pool_ ? pool_->threads() : n_threads; }

int MSystem::load(const std::string &config_path, int summation_at_load, bool analyse_at_load) {
    LineReader rd(config_path);
    if (!rd.f) { err = config_path + ": file not found"; return fail(EPH_EIO, err); }
    static const char *keys[] = {"Initial", "Extra", "Geopotentials", "Rings", "TDB", "Integrator", "Delta_t",
                                 "Cadence", "Max_Ephemeris_Length", "Combined_Delta_t_Max", "Combined_GM_Max_Parent",
                                 "Combined_GM_Max_Child", "Combined_GM_Max_Tiny", "Combined_Period_Max_Child"};
    enum { kInitial, kExtra, kGp, kRings, kTDB, kIntegrator, kDelta, kCadence, kMaxLen, kCDelta, kCParent, kCChild,
           kCTiny, kCPeriod };
    std::vector<std::string> vals(sizeof keys / sizeof keys[0]);
    std::string line;
    std::vector<char> k(1100), v(1100);
    while (rd.next(line)) {                               // mass_loader.cpp:120-158
        if (line.size() >= 1024) { err = "config line too long"; return fail(EPH_EIO, err); }
        if (sscanf(line.c_str(), "%s%s", k.data(), v.data()) != 2) { err = "config line length error: " + line; return fail(EPH_EIO, err); }
        size_t q = 0;
        while (q < vals.size() && strcmp(keys[q], k.data())) ++q;
        if (q == vals.size()) { err = std::string("unknown config parameter: ") + k.data(); return fail(EPH_EIO, err); }
        if (!vals[q].empty()) { err = std::string("duplicate config parameter: ") + k.data(); return fail(EPH_EIO, err); }
        vals[q] = v.data();
    }
    if (vals[kInitial].empty()) { err = "config has no Initial"; return fail(EPH_EIO, err); }
    const std::string dir = dir_of(config_path) + "/";
    int rc = load_files(dir + vals[kInitial], vals[kExtra].empty() ? "" : dir + vals[kExtra],
                        vals[kGp].empty() ? "" : dir + vals[kGp], vals[kRings].empty() ? "" : dir + vals[kRings],
                        summation_at_load, analyse_at_load);
    if (rc) return rc;
    if (!vals[kIntegrator].empty()) {                     // mass_loader.cpp:188-204
        static const char *names[] = {"CPU_RK12", "GPU_RK12", "COMBINED_RK12"};
        integrator_ = -1;
        for (int q = 0; q < 3; ++q)
            if (vals[kIntegrator] == names[q]) integrator_ = q;
        if (integrator_ < 0) { err = "invalid integrator: " + vals[kIntegrator]; return fail(EPH_EIO, err); }
    }
    auto load_param = [&](int key, double &param) {       // LOAD_CONFIG, mass_loader.cpp:206-211
        if (!vals[key].empty()) param = atof(vals[key].c_str());
    };
    double tdb = 0;
    load_param(kTDB, tdb);
    load_param(kDelta, delta_t_);
    load_param(kCadence, data_cadence);
    load_param(kMaxLen, max_ephm_length);
    if (integrator_ == COMBINED_RK12) {
        load_param(kCDelta, combined_delta_t);
        load_param(kCParent, GM_max_parent);
        load_param(kCChild, GM_max_child);
        load_param(kCTiny, GM_max_tiny);
        load_param(kCPeriod, Period_max_child);
    }
    t_eph_ = dd{tdb, 0};
    if (!(std::fabs(tdb) < 0x1p63) || (double)(int64_t)tdb != tdb) {   // mass_loader.cpp:226-229
        err = "TDB (initial ephemeris time, seconds) should be an integer";
        return fail(EPH_EIO, err);
    }
    for (auto &m : mlist) {                               // correct time variables to t = 0, mass_loader.cpp:231-237
        m.GM0 = m.GM - m.dGM * tdb;
        m.exJ2 = m.dJ2 * tdb;
        m.C_static.x.x -= m.exJ2 / 2;
        m.C_static.y.y -= m.exJ2 / 2;
        m.C_static.z.z += m.exJ2;
    }
    auto positive = [](double x) { return x > 0 && x != INFINITY; };   // mass_loader.cpp:14-16, 239-243
    if (!positive(delta_t_) || !positive(data_cadence) || !positive(max_ephm_length) || !positive(combined_delta_t)) {
        err = "Delta_t/Cadence/Max_Ephemeris_Length/Combined_Delta_t_Max should be finite positive real numbers";
        return fail(EPH_EIO, err);
    }
    host_changed();
    return EPH_OK;
}

int MSystem::load_files(const std::string &initial, const std::string &extra, const std::string &gp_dir,
                        const std::string &ring_dir, int summation_at_load, bool analyse_at_load) {
    mlist.clear();
    gps.clear();
    rgs.clear();
    diags.clear();
    blist.clear();
    tidal_children.clear();
    tidal_parent = -1;
    drop_device();
    reset_params();                                       // mass_loader.cpp:329
    LineReader rd(initial);
    if (!rd.f) { err = initial + ": file not found"; return fail(EPH_EIO, err); }
    std::string line;
    std::vector<char> sname(1100), sid(1100);
    while (rd.next(line)) {                               // mass_loader.cpp:351-471
        if (line.size() >= 1024) { err = initial + ": line too long"; return fail(EPH_EIO, err); }
        eph_mass m;
        memset(&m, 0, sizeof m);
        double j2, c21, c22, s21, s22, ra, dec, W, p, gprf, ringmf;
        double r[3], vv[3], x[3], z[3], w[3];
        int n = sscanf(line.c_str(),
                       "%[^\t]%s%lf%lf" "%lf%lf%lf%lf" "%lf%lf%lf%lf%lf" "%lf%lf%lf"
                       "%lf%lf%lf%lf%lf%lf" "%lf%lf%lf%lf" "%lf%lf%lf%lf%lf%lf%lf%lf%lf",
                       sname.data(), sid.data(), &m.GM, &m.R, &m.inertia, &m.k2, &m.k2r, &m.tide_delay,
                       &j2, &c21, &c22, &s21, &s22, &m.tide_delay_factor, &gprf, &ringmf,
                       &r[0], &r[1], &r[2], &vv[0], &vv[1], &vv[2], &ra, &dec, &W, &p,
                       &x[0], &x[1], &x[2], &z[0], &z[1], &z[2], &w[0], &w[1], &w[2]);
        if (!(n == 35 || n == 26)) { err = initial + ": line length error: " + line; return fail(EPH_EIO, err); }
        const size_t sl = strlen(sid.data());
        if (sl >= 8) { err = initial + ": sid too long: " + sid.data(); return fail(EPH_EIO, err); }
        memcpy(&m.sid, sid.data(), sl);
        for (const auto &o : mlist)
            if (o.sid == m.sid) { err = initial + ": duplicate sid " + sid.data(); return fail(EPH_EIO, err); }
        m.R *= 1000;
        m.GM0 = m.GM;
        m.gp = m.ring = -1;
        if (gprf != 0) {
            eph_geopotential g;
            int rc = load_gp((gp_dir.empty() ? std::string(".") : gp_dir) + "/" + sid.data() + ".txt", gprf, g, err);
            if (rc) return fail(rc, err);
            m.gp = (int64_t)gps.size();
            gps.push_back(g);
        }
        if (ringmf != 0) {
            eph_ring g;
            int rc = load_ring((ring_dir.empty() ? std::string(".") : ring_dir) + "/" + sid.data() + ".txt", m.GM, m.R,
                               ringmf, g, err);
            if (rc) return fail(rc, err);
            if (g.N > 0) { m.ring = (int64_t)rgs.size(); rgs.push_back(g); }
        }
        v3 X{x[0], x[1], x[2]}, Z{z[0], z[1], z[2]}, Wv{w[0], w[1], w[2]};
        if (n == 26) {                                    // mass_loader.cpp:441-451
            double th = (90 - dec) * kDegree, ph = ra * kDegree, c = std::sin(th);
            Z = v3{c * std::cos(ph), c * std::sin(ph), std::cos(th)};
            th = 90 * kDegree; ph = (ra + 90) * kDegree; c = std::sin(th);
            X = v3{c * std::cos(ph), c * std::sin(ph), std::cos(th)};
            X = X + rotation_matrix(Z, W * kDegree, HostSin{}) * X;
            const double sb = std::sin(-kObliquity), cb = std::cos(-kObliquity);
            double ny = X.y * cb - X.z * sb; X.z = X.z * cb + X.y * sb; X.y = ny;
            ny = Z.y * cb - Z.z * sb; Z.z = Z.z * cb + Z.y * sb; Z.y = ny;
            Wv = (2 * kPi / (p * 3600)) * Z;
        }
        ddm3 s{mkddv(X), mkddv(cross(Z, X)), mkddv(Z)};
        m.s = E(orthogonalized(s));
        m.r = E(mkddv(v3{r[0], r[1], r[2]}));
        m.v = E(mkddv(v3{vv[0], vv[1], vv[2]}));
        m.w = E(Wv);
        m.A = 3 * m.inertia / 2;
        m.R2 = m.R * m.R;
        m.sR2 = m.R < 0 ? m.R2 : 0;
        const m3 cs = from_harmonics(j2, c21, c22, s21, s22);
        m.C_static = E(cs);
        const m3 fmis = todbl(DM(m.s));
        const m3 GI = (-2.0 / 3 * m.R2) * subdiag(toworld(fmis, cs), m.A);
        m.GI = E(GI);
        m.GL = E(mkddv(GI * Wv));
        mlist.push_back(m);
    }
    if (mlist.empty()) { err = initial + ": no bodies"; return fail(EPH_EIO, err); }
    if (!extra.empty()) {                                 // mass_loader.cpp:475-538
        LineReader re(extra);
        if (!re.f) { err = extra + ": file not found"; return fail(EPH_EIO, err); }
        while (re.next(line)) {
            double param;
            if (sscanf(line.c_str(), "%s%s%lf", sid.data(), sname.data(), &param) != 3) {
                err = extra + ": line length error: " + line;
                return fail(EPH_EIO, err);
            }
            const int64_t mid = get_mid(sid.data());
            if (mid < 0) continue;
            eph_mass &m = mlist[(size_t)mid];
            if (!strcmp(sname.data(), "MassRate")) m.dGM = param;
            else if (!strcmp(sname.data(), "Luminosity")) m.lum = param;
            else if (!strcmp(sname.data(), "Receptance")) { m.recpt = param; m.rR2G_4c = m.recpt * m.R2 * (kG / (4 * kC)); }
            else if (!strcmp(sname.data(), "J2Rate")) m.dJ2 = param;
            else { err = extra + ": unknown parameter " + sname.data(); return fail(EPH_EIO, err); }
        }
    }
    return finish_load(summation_at_load, analyse_at_load);
}

int MSystem::finish_load(int summation_at_load, bool analyse_at_load) {
    const CpuSystem cs = cpu_view();
    cpu_deform(cs, pool());                               // msystem::deform, mass_loader.cpp:541
    std::string why;
    for (auto &m : mlist) {                               // mass_loader.cpp:544-551
        const m3 fmis = todbl(DM(m.s));
        const m3 cs_ = M(m.C_static);
        const m3 ncp = toworld(fmis, cs_);
        m.C_static = E(cs_ - (tolocal(fmis, M(m.C_potential)) - cs_));
        m.C_potential = E(ncp);
        std::string w;
        if (!mass_sanity(m, &w)) {
            char id[9] = {0};
            memcpy(id, &m.sid, 8);
            why += std::string("<") + id + ">: " + w;
        }
    }
    if (!why.empty()) { err = "sanity check failed: " + why; return fail(EPH_ESANITY, err); }
    int rc = cpu_accel_now(summation_at_load == EPH_SUMMATION_SERIAL ? -1 : 0);   // mass_loader.cpp:555
    if (rc) return rc;
    if (analyse_at_load) analyse(true);                   // mass_loader.cpp:558
    return EPH_OK;
}

int MSystem::assign(int64_t n, const eph_mass *masses, const eph_geopotential *g, int64_t n_gp, const eph_ring *r,
                    int64_t n_ring, dd t) {
    if (n <= 0 || !masses) { err = "no bodies"; return fail(EPH_EINVAL, err); }
    for (int64_t i = 0; i < n; ++i)
        if (masses[i].gp >= n_gp || masses[i].ring >= n_ring) { err = "model index out of range"; return fail(EPH_EINVAL, err); }
    drop_device();
    mlist.assign(masses, masses + n);
    gps.assign(g, g + (g ? n_gp : 0));
    rgs.assign(r, r + (r ? n_ring : 0));
    for (auto &m : mlist) {
        if (m.gp < 0) m.gp = -1;
        if (m.ring < 0) m.ring = -1;
    }
    t_eph_ = t;
    diags.clear();
    blist.clear();
    t_barycen = t_update = dd{NAN, NAN};
    tidal_children.clear();
    tidal_parent = -1;
    return EPH_OK;
}

int MSystem::set_masses(const eph_mass *m, int64_t n) {
    if (n != size() || !m) { err = "body count differs from the system's"; return fail(EPH_EINVAL, err); }
    for (int64_t i = 0; i < n; ++i)
        if (m[i].gp != mlist[(size_t)i].gp || m[i].ring != mlist[(size_t)i].ring) {
            err = "model indices differ from the system's";
            return fail(EPH_EINVAL, err);
        }
    std::copy(m, m + n, mlist.begin());
    host_changed();
    return EPH_OK;
}

CpuSystem MSystem::cpu_view() {
    CpuSystem cs;
    cs.m = mlist.data();
    cs.n = size();
    cs.gps = gps.data();
    cs.rings = rgs.data();
    cs.tidal_parent = tidal_parent;
    cs.tidal_matrix = tidal_matrix;
    cs.tidal_children = tidal_children.data();
    cs.n_tidal_children = (int64_t)tidal_children.size();
    return cs;
}

int MSystem::ensure_device() {
    if (!sys) {
        const int64_t n = size();
        const double t0[2] = {t_eph_.hi, t_eph_.lo};
        int rc = eph_system_create(1, &n, mlist.data(), gps.data(), (int64_t)gps.size(), rgs.data(),
                                   (int64_t)rgs.size(), t0, &sys);
        if (rc) { sys = nullptr; return rc; }
        gpu_stale = false;
    } else if (gpu_stale) {
        const double t0[2] = {t_eph_.hi, t_eph_.lo};
        int rc = eph_system_upload(sys, mlist.data(), t0);
        if (rc) return rc;
        gpu_stale = false;
    }
    return EPH_OK;
}

int MSystem::sync_down() {
    double t[2];
    int rc = eph_system_download(sys, mlist.data(), t);
    if (rc) return rc;
    t_eph_ = dd{t[0], t[1]};
    return EPH_OK;
}

int MSystem::cpu_accel_now(int parallel_option) {
    const CpuSystem cs = cpu_view();
    cpu_accel(cs, parallel_option, pool());
    host_changed();
    return EPH_OK;
}

int MSystem::accel(int parallel_option) {
    if (mlist.empty()) { err = "no system loaded"; return fail(EPH_EINVAL, err); }
    if (parallel_option <= 0) return cpu_accel_now(parallel_option);
    int rc = ensure_device();                             // msystem::Cuda_accel, mass.cpp:218-221
    if (rc) return rc;
    if ((rc = eph_accel(sys, EPH_SUMMATION_POOL))) return rc;
    return sync_down();
}

// collision and step-size diagnostics, mass.cpp:411-439
void MSystem::post_integrate(double dt) {
    diags.clear();
    const double TIMESTEP_THRESHOLD = 0.01;
    const double dt2 = dt * dt;
    for (int64_t i = 0; i < size(); ++i) {
        const eph_mass &m = mlist[(size_t)i];
        if (m.R > m.min_distance) diags.push_back(Diagnostic{i, 0, m.min_distance, m.R});
        if (m.max_influence * dt2 > TIMESTEP_THRESHOLD)
            diags.push_back(Diagnostic{i, 1, std::abs(dt), std::sqrt(TIMESTEP_THRESHOLD / m.max_influence)});
    }
}

int MSystem::integrate(double dt, int64_t n_step, int use_gpu) {
    if (mlist.empty()) { err = "no system loaded"; return fail(EPH_EINVAL, err); }
    if (!(dt == dt) || dt == 0) { err = "dt must be finite and non-zero"; return fail(EPH_EINVAL, err); }
    if (n_step == 0) return EPH_OK;
    if (use_gpu != CPU_RK12 && use_gpu != GPU_RK12) { err = "USE_GPU must be 0 or 1"; return fail(EPH_EINVAL, err); }
    const double dts = n_step < 0 ? -dt : dt;
    if (use_gpu == CPU_RK12) {
        const CpuSystem cs = cpu_view();
        cpu_integrate(cs, t_eph_, dt, n_step, pool());
        host_changed();
    } else {
        int rc = ensure_device();
        if (rc) return rc;
        if ((rc = eph_integrate(sys, dt, n_step))) return rc;
        if ((rc = sync_down())) return rc;
    }
    post_integrate(dts);
    return EPH_OK;
}

void MSystem::update(double t) {                          // msystem::update(t), mass.cpp:28-48
    const CpuSystem cs = cpu_view();
    cpu_update(cs, t);
    host_changed();
}

int MSystem::scale(int64_t mid, double factor) {         // mass::scale, mass.cpp:8-26
    if (mid < 0 || mid >= size()) { err = "body index out of range"; return fail(EPH_EINVAL, err); }
    eph_mass &m = mlist[(size_t)mid];
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
    host_changed();
    return EPH_OK;
}

int MSystem::scale_geopotential(int64_t mid, double factor) {   // mass_loader.cpp:1008-1021
    if (mid < 0 || mid >= size()) { err = "body index out of range"; return fail(EPH_EINVAL, err); }
    eph_mass &mi = mlist[(size_t)mid];
    if (mi.gp >= 0 && factor != 0) {                      // geopotential::copy(gp, factor), geopotential.cpp:154-166
        eph_geopotential g = gps[(size_t)mi.gp];
        if (factor != 1) {
            const int64_t nt = g.Nz * (g.Nz + 2) - 3;
            for (int64_t q = 0; q < nt; ++q) g.c[q] *= factor;
        }
        mi.gp = (int64_t)gps.size();
        gps.push_back(g);
    } else
        mi.gp = -1;
    mi.C_potential = E(factor * M(mi.C_potential));
    mi.C_static = E(factor * M(mi.C_static));
    mi.k2 *= factor;
    mi.k2r *= factor;
    mi.exJ2 *= factor;
    mi.dJ2 *= factor;
    const m3 GI = (-2.0 / 3 * mi.R2) * subdiag(M(mi.C_potential), mi.A);
    mi.GI = E(GI);
    mi.GL = E(mkddv(GI * V(mi.w)));
    drop_device();
    return EPH_OK;
}

int MSystem::scale_ring(int64_t mid, double factor) {    // mass_loader.cpp:1022-1026
    if (mid < 0 || mid >= size()) { err = "body index out of range"; return fail(EPH_EINVAL, err); }
    eph_mass &mi = mlist[(size_t)mid];
    if (mi.ring >= 0 && factor != 0) {                    // ring::copy(rp, factor), ring.cpp:118-134
        eph_ring r = rgs[(size_t)mi.ring];
        if (factor != 1) {
            const double host_ratio = 1 - r.GM_ratio;
            const double new_ratio = r.GM_ratio * factor;
            r.GM_ratio = new_ratio / (host_ratio + new_ratio);
            r.A *= factor;
            r.J2 *= factor;
            r.GL_R2 *= factor;
            for (int64_t q = 0; q < r.N; ++q) r.GsR2[q] *= factor;
        }
        mi.ring = (int64_t)rgs.size();
        rgs.push_back(r);
    } else
        mi.ring = -1;
    drop_device();
    return EPH_OK;
}

int MSystem::set_tidal(int64_t parent, const m3 &matrix, const std::vector<int64_t> &children) {
    for (int64_t c : children)
        if (c < 0 || c >= size() || c == parent) { err = "tidal child index out of range"; return fail(EPH_EINVAL, err); }
    if (!children.empty() && (parent < 0 || parent >= size())) { err = "tidal parent index out of range"; return fail(EPH_EINVAL, err); }
    tidal_parent = parent;
    tidal_matrix = matrix;
    tidal_children = children;
    return EPH_OK;
}

int64_t MSystem::get_mid(const char *ssid) const {      // mass_loader.cpp:901-908
    uint64_t key = 0;
    const size_t sl = strlen(ssid);
    if (sl >= 8) return -1;
    memcpy(&key, ssid, sl);
    return get_mid(key);
}

int64_t MSystem::get_mid(uint64_t sid) const {
    for (size_t i = 0; i < mlist.size(); ++i)
        if (mlist[i].sid == sid) return (int64_t)i;
    return -1;
}

int64_t MSystem::ephemeris_time() const {                // dfloat_t operator I, dfloat_t.h:62-67
    const dd t = t_eph_;
    int64_t ia = (int64_t)std::llround(t.hi + t.lo);
    const dd di{(double)ia, (double)(ia - (int64_t)(double)ia)};
    return ia + (int64_t)std::floor(todbl(sub(t, di)));
}

}  // namespace eph

// ---------------------------------------------------------------------------
// C ABI of the host class
// ---------------------------------------------------------------------------
using eph::MSystem;
static MSystem *H(eph_msystem *h) { return reinterpret_cast<MSystem *>(h); }
static const MSystem *H(const eph_msystem *h) { return reinterpret_cast<const MSystem *>(h); }

extern "C" {

int eph_msystem_load_ex(const char *config_path, int summation_at_load, int analyse, eph_msystem **out) {
    if (!config_path || !out) return eph::fail(EPH_EINVAL, "null argument");
    auto *m = new MSystem();
    int rc = m->load(config_path, summation_at_load, analyse != 0);
    if (rc) { delete m; return rc; }
    *out = reinterpret_cast<eph_msystem *>(m);
    return EPH_OK;
}

int eph_msystem_load(const char *config_path, int summation_at_load, eph_msystem **out) {
    return eph_msystem_load_ex(config_path, summation_at_load, 1, out);
}

int eph_msystem_create(int64_t n, const eph_mass *masses, const eph_geopotential *gps, int64_t n_gp,
                       const eph_ring *rings, int64_t n_ring, const double *t_eph, eph_msystem **out) {
    if (!masses || !out || n <= 0) return eph::fail(EPH_EINVAL, "null or empty argument");
    auto *m = new MSystem();
    int rc = m->assign(n, masses, gps, n_gp, rings, n_ring, t_eph ? eph::dd{t_eph[0], t_eph[1]} : eph::dd{0, 0});
    if (rc) { delete m; return rc; }
    *out = reinterpret_cast<eph_msystem *>(m);
    return EPH_OK;
}

int eph_msystem_load_checkpoint(const char *path, eph_msystem **out) {
    if (!path || !out) return eph::fail(EPH_EINVAL, "null argument");
    auto *m = new MSystem();
    int rc = m->load_checkpoint(std::string(path));
    if (rc) { delete m; return rc; }
    *out = reinterpret_cast<eph_msystem *>(m);
    return EPH_OK;
}

int eph_msystem_save_checkpoint(const eph_msystem *h, const char *path) {
    if (!h || !path) return eph::fail(EPH_EINVAL, "null argument");
    return H(h)->save_checkpoint(std::string(path));
}

int eph_msystem_destroy(eph_msystem *h) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    delete H(h);
    return EPH_OK;
}

int64_t eph_msystem_size(const eph_msystem *h) { return h ? H(h)->size() : -1; }

int eph_msystem_masses(const eph_msystem *h, eph_mass *out, int64_t cap) {
    if (!h || !out) return eph::fail(EPH_EINVAL, "null argument");
    if (cap < H(h)->size()) return eph::fail(EPH_EINVAL, "output capacity too small");
    for (int64_t i = 0; i < H(h)->size(); ++i) out[i] = (*H(h))[i];
    return EPH_OK;
}

int eph_msystem_set_masses(eph_msystem *h, const eph_mass *masses, int64_t n) {
    if (!h || !masses) return eph::fail(EPH_EINVAL, "null argument");
    return H(h)->set_masses(masses, n);
}

int eph_msystem_models(const eph_msystem *h, eph_geopotential *gps, int64_t gcap, eph_ring *rings, int64_t rcap,
                       int64_t *n_gp, int64_t *n_ring) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    const auto &g = H(h)->geopotentials();
    const auto &r = H(h)->rings();
    if (n_gp) *n_gp = (int64_t)g.size();
    if (n_ring) *n_ring = (int64_t)r.size();
    for (int64_t i = 0; gps && i < (int64_t)g.size() && i < gcap; ++i) gps[i] = g[(size_t)i];
    for (int64_t i = 0; rings && i < (int64_t)r.size() && i < rcap; ++i) rings[i] = r[(size_t)i];
    return EPH_OK;
}

int eph_msystem_time(const eph_msystem *h, double *t_eph) {
    if (!h || !t_eph) return eph::fail(EPH_EINVAL, "null argument");
    t_eph[0] = H(h)->t_eph().hi;
    t_eph[1] = H(h)->t_eph().lo;
    return EPH_OK;
}

int eph_msystem_set_time(eph_msystem *h, const double *t_eph) {
    if (!h || !t_eph) return eph::fail(EPH_EINVAL, "null argument");
    H(h)->set_t_eph(eph::dd{t_eph[0], t_eph[1]});
    return EPH_OK;
}

int eph_msystem_params(const eph_msystem *h, double *params8, int64_t *integrator) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    if (params8) H(h)->params(params8);
    if (integrator) *integrator = H(h)->integrator();
    return EPH_OK;
}

int eph_msystem_accel(eph_msystem *h, int parallel_option) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    return H(h)->accel(parallel_option);
}

int eph_msystem_integrate(eph_msystem *h, double dt, int64_t n_step, int use_gpu) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    return H(h)->integrate(dt, n_step, use_gpu);
}

int eph_msystem_combined_integrate(eph_msystem *h, double dt, int64_t n_combine, int64_t n_step, int use_gpu) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    return H(h)->combined_integrate(dt, n_combine, n_step, use_gpu);
}

int eph_msystem_update(eph_msystem *h, double t) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    H(h)->update(t);
    return EPH_OK;
}

int eph_msystem_analyse(eph_msystem *h, int reconstruct, double *t_update) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    const eph::dd t = H(h)->analyse(reconstruct != 0);
    if (t_update) { t_update[0] = t.hi; t_update[1] = t.lo; }
    if (!H(h)->analysis_ok())
        return eph::fail(EPH_EINVAL, "orbit-structure analysis did not settle; the previous structure is kept");
    return EPH_OK;
}

int64_t eph_msystem_barycens(const eph_msystem *h, int64_t *ids, double *gm_sys, int64_t *children, int64_t cap,
                             int64_t child_cap) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    const auto &bl = H(h)->barycens();
    int64_t nc = 0;
    for (int64_t i = 0; i < (int64_t)bl.size() && i < cap; ++i) {
        const eph::Barycen &b = bl[(size_t)i];
        if (ids) {
            int64_t *o = ids + 7 * i;
            o[0] = b.pid; o[1] = b.hid; o[2] = b.gid; o[3] = b.tid; o[4] = b.mid;
            o[5] = (int64_t)b.children.size();
            o[6] = nc;
        }
        if (gm_sys) { gm_sys[2 * i] = b.GM_sys.hi; gm_sys[2 * i + 1] = b.GM_sys.lo; }
        for (int64_t c : b.children) {
            if (children && nc < child_cap) children[nc] = c;
            ++nc;
        }
    }
    return (int64_t)bl.size();
}

int eph_msystem_scale(eph_msystem *h, int64_t mid, double factor) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    return H(h)->scale(mid, factor);
}

int eph_msystem_scale_geopotential(eph_msystem *h, int64_t mid, double factor) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    return H(h)->scale_geopotential(mid, factor);
}

int eph_msystem_scale_ring(eph_msystem *h, int64_t mid, double factor) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    return H(h)->scale_ring(mid, factor);
}

int eph_msystem_set_tidal(eph_msystem *h, int64_t parent, const double *matrix9, const int64_t *children,
                          int64_t n_children) {
    if (!h || (n_children > 0 && (!children || !matrix9))) return eph::fail(EPH_EINVAL, "null argument");
    eph::m3 m{};
    if (matrix9) m = eph::m3{eph::v3{matrix9[0], matrix9[1], matrix9[2]}, eph::v3{matrix9[3], matrix9[4], matrix9[5]},
                             eph::v3{matrix9[6], matrix9[7], matrix9[8]}};
    std::vector<int64_t> ch(children, children + (n_children > 0 ? n_children : 0));
    return H(h)->set_tidal(parent, m, ch);
}

int eph_msystem_set_threads(eph_msystem *h, int n_threads) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    return H(h)->set_threads(n_threads);
}

int64_t eph_msystem_get_mid(const eph_msystem *h, const char *ssid) {
    if (!h || !ssid) return eph::fail(EPH_EINVAL, "null argument");
    return H(h)->get_mid(ssid);
}

int eph_msystem_diagnostics(const eph_msystem *h, int64_t *body, int *kind, double *value, double *limit, int64_t cap) {
    if (!h) return eph::fail(EPH_EINVAL, "null handle");
    const auto &d = H(h)->diagnostics();
    for (int64_t i = 0; i < (int64_t)d.size() && i < cap; ++i) {
        if (body) body[i] = d[(size_t)i].body;
        if (kind) kind[i] = d[(size_t)i].kind;
        if (value) value[i] = d[(size_t)i].value;
        if (limit) limit[i] = d[(size_t)i].limit;
    }
    return (int)d.size();
}

}  // extern "C"
