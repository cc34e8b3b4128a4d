// C++ host class mirroring the reference's msystem operator interface
// (src/physics/mass.h:202-358): load a system from the text inputs or a
// checkpoint, accel(parallel_option), integrate(dt, n_step, USE_GPU),
// combined_integrate, the orbit-structure analysis, element access and the
// post-integration diagnostics.  The host copy of the bodies is
// authoritative; the GPU back end (the C ABI of include/ephemeris_hip.h)
// keeps a resident copy that is refreshed only when the host copy changed,
// and the CPU back end (eph_cpu.hpp) works on the host copy directly.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ephemeris_hip.h"
#include "eph_math.hpp"

namespace eph {

class WorkerPool;
struct CpuSystem;

struct Diagnostic {
    int64_t body;          // index in the system
    int kind;              // 0 = collision (another body inside R), 1 = time step too large
    double value, limit;   // min distance vs R, or dt vs recommended dt
};

// barycen of the reference (mass.h:11-45): a body or the barycentre of a
// host-guest pair, with its place in the orbit hierarchy
struct Barycen {
    ddv3 r{}, v{};
    dd GM{};
    ddv3 r_sys{}, v_sys{};
    dd GM_sys{};
    int64_t pid = -1, hid = -1, gid = -1, tid = -1, mid = -1;
    std::vector<int64_t> children;
};

class MSystem {
public:
    // integrator codes of the reference (mass.h:207-211)
    enum { CPU_RK12 = 0, GPU_RK12 = 1, COMBINED_RK12 = 2 };

    MSystem();
    ~MSystem();
    MSystem(const MSystem &) = delete;
    MSystem &operator=(const MSystem &) = delete;

    // msystem::load(fconfig) (src/modules/mass_loader.cpp:67-294): the config
    // names the Initial, Extra, Geopotentials and Rings inputs and the
    // integration parameters.  The initial force evaluation runs on the CPU
    // back end in the order `summation_at_load` selects (the reference loads
    // from a thread without a pool: EPH_SUMMATION_SERIAL).
    // analyse = false skips the orbit-structure analysis of the loader (for
    // callers that never use the structure: ensemble members, particle fields)
    int load(const std::string &config_path, int summation_at_load = EPH_SUMMATION_SERIAL, bool analyse = true);
    // msystem::load(fbase, fext, gppath, ringpath) (mass_loader.cpp:324-561)
    int load_files(const std::string &initial, const std::string &extra, const std::string &gp_dir,
                   const std::string &ring_dir, int summation_at_load = EPH_SUMMATION_SERIAL, bool analyse = true);
    // a system from bodies and models in memory (model indices into gps/rings)
    int assign(int64_t n, const eph_mass *masses, const eph_geopotential *gps, int64_t n_gp, const eph_ring *rings,
               int64_t n_ring, dd t);
    // msystem::load_checkpoint / save_checkpoint (mass_loader.cpp:614-759),
    // the reference's binary layout
    int load_checkpoint(const std::string &path);
    int save_checkpoint(const std::string &path) const;
    int save_checkpoint(std::vector<char> &out) const;
    int load_checkpoint(const char *data, size_t size);

    // msystem::accel (mass.cpp:216-397): 1 = GPU, 0 = CPU thread pool (the
    // pool order when min(N, ceil(N^2/2304)) > 1), -1 = CPU single thread
    int accel(int parallel_option = 0);
    // msystem::integrate (mass.cpp:399-442): USE_GPU 0 = CPU RK12, 1 = GPU RK12
    int integrate(double dt, int64_t n_step, int use_gpu = GPU_RK12);
    // msystem::combined_integrate (src/integrators/mass_combined.cpp:40-406)
    int combined_integrate(double dt, int64_t n_combine, int64_t n_step, int use_gpu = 1);
    // msystem::update(t) (mass.cpp:28-48): GM and exJ2 of every body at t
    void update(double t);
    // msystem::analyse (src/physics/barycen_analyzer.cpp:12-310); returns
    // t_update.  analysis_ok() is false when the last analysis did not settle
    // within a bounded number of passes (the reference would loop for ever)
    // and the previous structure was kept.
    dd analyse(bool reconstruct = false);
    bool analysis_ok() const { return analysis_settled; }
    // mass::scale and msystem::scale_geopotential / scale_ring
    // (mass.cpp:8-26, mass_loader.cpp:1005-1026); accel() must follow
    int scale(int64_t mid, double factor);
    int scale_geopotential(int64_t mid, double factor);
    int scale_ring(int64_t mid, double factor);
    // tidal-matrix correction applied by the CPU accel (mass.cpp:376-387)
    int set_tidal(int64_t parent, const m3 &matrix, const std::vector<int64_t> &children);
    // worker threads of the CPU back end (0 = hardware concurrency)
    int set_threads(int n);
    int threads() const;

    int64_t size() const { return (int64_t)mlist.size(); }
    const eph_mass &operator[](int64_t i) const { return mlist[(size_t)i]; }
    eph_mass &at(int64_t i) { host_changed(); return mlist[(size_t)i]; }
    int set_masses(const eph_mass *m, int64_t n);
    int64_t get_mid(const char *ssid) const;
    int64_t get_mid(uint64_t sid) const;
    int64_t ephemeris_time() const;                 // floor of t_eph (dfloat_t.h:62-67)
    dd t_eph() const { return t_eph_; }
    void set_t_eph(dd t) { t_eph_ = t; host_changed(); }
    double delta_t() const { return delta_t_; }
    int64_t integrator() const { return integrator_; }
    // integration parameters of the config (mass.h:214-232), in the order of
    // the checkpoint header: delta_t, data_cadence, max_ephm_length,
    // combined_delta_t, GM_max_child, GM_max_parent, GM_max_tiny, Period_max_child
    void params(double *out8) const;
    const std::vector<eph_geopotential> &geopotentials() const { return gps; }
    const std::vector<eph_ring> &rings() const { return rgs; }
    const std::vector<Barycen> &barycens() const { return blist; }
    const std::vector<Diagnostic> &diagnostics() const { return diags; }
    const std::string &error() const { return err; }

private:
    int finish_load(int summation_at_load, bool analyse);
    CpuSystem cpu_view();
    int ensure_device();
    int sync_down();
    void host_changed() { gpu_stale = true; }
    void drop_device();
    WorkerPool *pool();
    int cpu_accel_now(int parallel_option);
    void post_integrate(double dt);
    void reset_params();
    void copy_params(const MSystem &o);

    std::vector<eph_mass> mlist;
    std::vector<eph_geopotential> gps;
    std::vector<eph_ring> rgs;
    std::vector<Diagnostic> diags;
    dd t_eph_{0, 0};
    double delta_t_ = 300, data_cadence = 7200, max_ephm_length = 631152000, combined_delta_t = 28800;
    double GM_max_child = 1E13, GM_max_parent = 2E17, GM_max_tiny = 14E11, Period_max_child = 0;
    int64_t integrator_ = COMBINED_RK12;
    // orbit structure (barycen_analyzer.cpp)
    std::vector<Barycen> blist;
    dd t_barycen{NAN, NAN}, t_update{NAN, NAN};
    bool analysis_settled = true;
    // tidal-matrix correction of a combined-integration subsystem
    int64_t tidal_parent = -1;
    m3 tidal_matrix{};
    std::vector<int64_t> tidal_children;
    // back ends
    int n_threads = 0;
    std::unique_ptr<WorkerPool> pool_;
    eph_system *sys = nullptr;
    bool gpu_stale = true;
    std::string err;
};

// sanity checks of one body (mass::sanity, mass_loader.cpp:18-65); appends
// the failures to `why` and returns true when the body is valid
bool mass_sanity(const eph_mass &m, std::string *why);

// bytes of class mass written per body by the reference's checkpoint: the
// fields from r up to the temporaries (mass.h:67-134; Egrad onwards is not
// stored)
constexpr size_t kCheckpointMassBytes = offsetof(eph_mass, Egrad);

}  // namespace eph
