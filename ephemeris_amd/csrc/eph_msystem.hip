// Text-format loader and the msystem-style host class (see eph_msystem.hpp).
#include "eph_msystem.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

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

MSystem::~MSystem() {
    if (sys) eph_system_destroy(sys);
}

int MSystem::load(const std::string &config_path, int summation_at_load) {
    LineReader rd(config_path);
    if (!rd.f) { err = config_path + ": file not found"; return fail(EPH_EIO, err); }
    static const char *keys[] = {"Initial", "Extra", "Geopotentials", "Rings", "TDB", "Integrator", "Delta_t",
                                 "Cadence", "Max_Ephemeris_Length", "Combined_Delta_t_Max", "Combined_GM_Max_Parent",
                                 "Combined_GM_Max_Child", "Combined_GM_Max_Tiny", "Combined_Period_Max_Child"};
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
    if (vals[0].empty()) { err = "config has no Initial"; return fail(EPH_EIO, err); }
    const std::string dir = dir_of(config_path) + "/";
    int rc = load_files(dir + vals[0], vals[1].empty() ? "" : dir + vals[1], vals[2].empty() ? "" : dir + vals[2],
                        vals[3].empty() ? "" : dir + vals[3], summation_at_load);
    if (rc) return rc;
    if (!vals[4].empty()) { t_eph[0] = (double)atoll(vals[4].c_str()); t_eph[1] = 0; }
    if (!vals[6].empty()) delta_t_ = atof(vals[6].c_str());
    if (!vals[4].empty()) return eph_system_upload(sys, mlist.data(), t_eph);
    return EPH_OK;
}

int MSystem::load_files(const std::string &initial, const std::string &extra, const std::string &gp_dir,
                        const std::string &ring_dir, int summation_at_load) {
    mlist.clear();
    gps.clear();
    rgs.clear();
    diags.clear();
    if (sys) { eph_system_destroy(sys); sys = nullptr; }
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
    return finish_load(summation_at_load);
}

int MSystem::finish_load(int summation_at_load) {
    const int64_t n = size();
    const double t0[2] = {t_eph[0], t_eph[1]};
    int rc = eph_system_create(1, &n, mlist.data(), gps.data(), (int64_t)gps.size(), rgs.data(), (int64_t)rgs.size(),
                               t0, &sys);
    if (rc) return rc;
    if ((rc = eph_deform(sys))) return rc;                // msystem::deform, mass_loader.cpp:541
    if ((rc = eph_system_download(sys, mlist.data(), nullptr))) return rc;
    std::string why;
    for (auto &m : mlist) {                               // mass_loader.cpp:544-551
        const m3 fmis = todbl(DM(m.s));
        const m3 cs = M(m.C_static);
        const m3 ncp = toworld(fmis, cs);
        m.C_static = E(cs - (tolocal(fmis, M(m.C_potential)) - cs));
        m.C_potential = E(ncp);
        std::string w;
        if (!mass_sanity(m, &w)) {
            char id[9] = {0};
            memcpy(id, &m.sid, 8);
            why += std::string("<") + id + ">: " + w;
        }
    }
    if (!why.empty()) { err = "sanity check failed: " + why; return fail(EPH_ESANITY, err); }
    if ((rc = eph_system_upload(sys, mlist.data(), t0))) return rc;
    return accel(summation_at_load == EPH_SUMMATION_SERIAL ? -1 : 0);
}

int MSystem::sync_down() {
    return eph_system_download(sys, mlist.data(), t_eph);
}

int MSystem::accel(int parallel_option) {
    if (!sys) { err = "no system loaded"; return fail(EPH_EINVAL, err); }
    int rc = eph_accel(sys, parallel_option < 0 ? EPH_SUMMATION_SERIAL : EPH_SUMMATION_POOL);
    if (rc) return rc;
    return sync_down();
}

int MSystem::integrate(double dt, int64_t n_step, int use_gpu) {
    if (!sys) { err = "no system loaded"; return fail(EPH_EINVAL, err); }
    if (use_gpu != GPU_RK12) {
        err = "only the GPU back end (USE_GPU=1) is provided by this library";
        return fail(EPH_ENOTSUP, err);
    }
    if (n_step == 0) return EPH_OK;
    int rc = eph_integrate(sys, dt, n_step);
    if (rc) return rc;
    if ((rc = sync_down())) return rc;
    // collision and step-size diagnostics, mass.cpp:411-439
    diags.clear();
    const double TIMESTEP_THRESHOLD = 0.01;
    const double dt2 = dt * dt;
    for (int64_t i = 0; i < size(); ++i) {
        const eph_mass &m = mlist[(size_t)i];
        if (m.R > m.min_distance) diags.push_back(Diagnostic{i, 0, m.min_distance, m.R});
        if (m.max_influence * dt2 > TIMESTEP_THRESHOLD)
            diags.push_back(Diagnostic{i, 1, std::abs(dt), std::sqrt(TIMESTEP_THRESHOLD / m.max_influence)});
    }
    return EPH_OK;
}

int64_t MSystem::get_mid(const char *ssid) const {
    uint64_t key = 0;
    const size_t sl = strlen(ssid);
    if (sl >= 8) return -1;
    memcpy(&key, ssid, sl);
    for (size_t i = 0; i < mlist.size(); ++i)
        if (mlist[i].sid == key) return (int64_t)i;
    return -1;
}

int64_t MSystem::ephemeris_time() const {                // dfloat_t operator I, dfloat_t.h:62-67
    const dd t{t_eph[0], t_eph[1]};
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

int eph_msystem_load(const char *config_path, int summation_at_load, eph_msystem **out) {
    if (!config_path || !out) return eph::fail(EPH_EINVAL, "null argument");
    auto *m = new MSystem();
    int rc = m->load(config_path, summation_at_load);
    if (rc) { delete m; return rc; }
    *out = reinterpret_cast<eph_msystem *>(m);
    return EPH_OK;
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
    t_eph[0] = H(h)->t_eph_hi();
    t_eph[1] = H(h)->t_eph_lo();
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
