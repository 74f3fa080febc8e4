// hip_solver.hpp -- the HIP backend classes, the drop-ins beside the reference's
// FastHLLMUSCLBathyOpenMPSolver / FastHLLMUSCLBathyCUDASolver
// (include/backend/cuda/cuda_solver.hpp:109-230) and its modular SerialSolver
// (include/backend/serial/serial_solver.hpp:101-300): constructed from a SimulationConfig,
// run() advances the configured number of steps and writes the snapshots.
//
// Host-only C++ (g++): the device work goes through the C ABI of libswe_hip.so
// (include/swe_hip.h), so this file includes no HIP header.  Every failure becomes a
// std::runtime_error carrying the backend's status string and detail message.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "config.hpp"
#include "netcdf_classic.hpp"
#include "sanity.hpp"
#include "setup.hpp"

struct swe_hip_ctx;

struct HIPSolverTimingStats {
    double cfl_check = 0.0;
    double host_to_device = 0.0;
    double device_to_host = 0.0;
    double steps = 0.0;  // the fused stage launches (boundary, rhs, combine, floor)
    double output = 0.0;
    double sanity_checks = 0.0;
    std::size_t time_steps = 0;
    void print_summary() const;
};

// The scheme family a solver runs (swe_scheme / swe_recon / swe_riemann of swe_hip.h).
struct HIPScheme {
    int scheme = 0;  // SWE_SCHEME_OPTIMIZED
    int recon = 0;   // SWE_RECON_MUSCL
    int riemann = 0; // SWE_RIEMANN_HLL
};

// Shared body of the HIP solver classes below: grid, host state, output writers, sanity
// checks and the step loop over the C ABI.
class HIPSolverBase {
  public:
    ~HIPSolverBase();
    HIPSolverBase(const HIPSolverBase &) = delete;
    HIPSolverBase &operator=(const HIPSolverBase &) = delete;

    void run();

    // The state on the host after run() (reference layout, ghosts filled).
    const std::vector<double> &h() const { return h_; }
    const std::vector<double> &hu() const { return hu_; }
    const std::vector<double> &hv() const { return hv_; }

  protected:
    HIPSolverBase(const SimulationConfig &cfg, HIPScheme scheme);

  private:
    void check(int rc, const char *what) const;
    void download();
    void run_sanity_checks(double time, std::size_t step);
    double serial_stable_dt() const;

    const SimulationConfig cfg_;
    const HIPScheme scheme_;
    const std::string riemann_name_, recon_name_, bathy_name_;
    HostGrid grid_;
    double dt_;
    std::size_t steps_, save_every_;
    std::vector<double> h_, hu_, hv_, B_;
    NetCDFWriter writer_;
    std::vector<std::unique_ptr<SanityCheck>> sanity_checks_;
    std::unique_ptr<SanityCheckNetCDFWriter> sanity_writer_;
    swe_hip_ctx *ctx_ = nullptr;
    HIPSolverTimingStats timing_;
};

// The drop-in beside FastHLLMUSCLBathyCUDASolver: the optimized MUSCL + HLL + hydrostatic
// reconstruction scheme whatever [solver] says (solver_assembly.hpp:124-160 warns).
class FastHLLMUSCLBathyHIPSolver : public HIPSolverBase {
  public:
    explicit FastHLLMUSCLBathyHIPSolver(const SimulationConfig &cfg)
        : HIPSolverBase(cfg, HIPScheme{}) {}
};

// The reference's modular SerialSolver<ReflectingWalls, Recon, Riemann, SSPRK3, Bathymetry>
// (include/backend/serial/serial_solver.hpp) on the GPU: reconstruction and Riemann solver
// from [solver], the hydrostatic flux assembly with Flat / GaussHill bathymetry (HLL), the
// plain one with bathymetry None; the serial solver's CFL-check formula
// (time_utils.hpp:13-36) and metadata names.
class SerialHIPSolver : public HIPSolverBase {
  public:
    explicit SerialHIPSolver(const SimulationConfig &cfg);
};

// solver_assembly.hpp:124-160: the fixed scheme of the optimized backends, and the config
// settings it ignores (printed to stderr).
void warn_if_ignored_optimized_settings(const SimulationConfig &cfg, const std::string &backend);

// time_utils.hpp:40-71
double estimate_eta_seconds(std::size_t completed, std::size_t total, double elapsed);
std::string format_duration(double seconds);
