// C++ host class mirroring the reference's msystem operator interface for the
// hot path (src/physics/mass.h:202-358): load a system from the text inputs,
// accel(parallel_option), integrate(dt, n_step, USE_GPU), element access and
// the post-integration diagnostics.  All compute goes through the C ABI of
// include/ephemeris_hip.h; there is no CPU back end.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ephemeris_hip.h"

namespace eph {

struct Diagnostic {
    int64_t body;          // index in the system
    int kind;              // 0 = collision (another body inside R), 1 = time step too large
    double value, limit;   // min distance vs R, or dt vs recommended dt
};

class MSystem {
public:
    // CPU_RK12 / GPU_RK12 of the reference's integrate(); only the GPU is provided
    enum { CPU_RK12 = 0, GPU_RK12 = 1 };

    MSystem() = default;
    ~MSystem();
    MSystem(const MSystem &) = delete;
    MSystem &operator=(const MSystem &) = delete;

    // msystem::load(fconfig) (src/modules/mass_loader.cpp:67-294): the config
    // names the Initial, Extra, Geopotentials and Rings inputs.  The initial
    // force evaluation uses `summation_at_load` (the reference loads without a
    // thread pool: EPH_SUMMATION_SERIAL).
    int load(const std::string &config_path, int summation_at_load = EPH_SUMMATION_SERIAL);
    // msystem::load(fbase, fext, gppath, ringpath) (mass_loader.cpp:324-561)
    int load_files(const std::string &initial, const std::string &extra, const std::string &gp_dir,
                   const std::string &ring_dir, int summation_at_load = EPH_SUMMATION_SERIAL);

    // parallel_option: 1 = GPU with the thread-pool summation order of the
    // reference CPU path, 0 likewise, -1 = single-thread summation order.
    int accel(int parallel_option = 0);
    // USE_GPU must be GPU_RK12; n_step < 0 integrates backwards.
    int integrate(double dt, int64_t n_step, int use_gpu = GPU_RK12);

    int64_t size() const { return (int64_t)mlist.size(); }
    const eph_mass &operator[](int64_t i) const { return mlist[(size_t)i]; }
    int64_t get_mid(const char *ssid) const;
    int64_t ephemeris_time() const;                 // floor of t_eph (dfloat_t.h:62-67)
    double t_eph_hi() const { return t_eph[0]; }
    double t_eph_lo() const { return t_eph[1]; }
    double delta_t() const { return delta_t_; }
    const std::vector<eph_geopotential> &geopotentials() const { return gps; }
    const std::vector<eph_ring> &rings() const { return rgs; }
    const std::vector<Diagnostic> &diagnostics() const { return diags; }
    const std::string &error() const { return err; }

private:
    int finish_load(int summation_at_load);
    int sync_down();

    std::vector<eph_mass> mlist;
    std::vector<eph_geopotential> gps;
    std::vector<eph_ring> rgs;
    std::vector<Diagnostic> diags;
    double t_eph[2] = {0, 0};
    double delta_t_ = 300;
    eph_system *sys = nullptr;
    std::string err;
};

// sanity checks of one body (mass::sanity, mass_loader.cpp:18-65); appends
// the failures to `why` and returns true when the body is valid
bool mass_sanity(const eph_mass &m, std::string *why);

}  // namespace eph
