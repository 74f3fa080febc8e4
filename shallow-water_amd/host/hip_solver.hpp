// hip_solver.hpp -- the HIP backend class, the drop-in beside the reference's
// FastHLLMUSCLBathyOpenMPSolver / FastHLLMUSCLBathyCUDASolver
// (include/backend/cuda/cuda_solver.hpp:109-230): constructed from a SimulationConfig, run()
// advances the configured number of steps and writes the snapshots.
//
// Host-only C++ (g++): the device work goes through the C ABI of libswe_hip.so
// (include/swe_hip.h), so this file includes no HIP header.  Every failure becomes a
// std::runtime_error carrying the backend's status string and detail message.
#pragma once

#include <memory>
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

class FastHLLMUSCLBathyHIPSolver {
  public:
    explicit FastHLLMUSCLBathyHIPSolver(const SimulationConfig &cfg);
    ~FastHLLMUSCLBathyHIPSolver();
    FastHLLMUSCLBathyHIPSolver(const FastHLLMUSCLBathyHIPSolver &) = delete;
    FastHLLMUSCLBathyHIPSolver &operator=(const FastHLLMUSCLBathyHIPSolver &) = delete;

    void run();

    // The state on the host after run() (reference layout, ghosts filled).
    const std::vector<double> &h() const { return h_; }
    const std::vector<double> &hu() const { return hu_; }
    const std::vector<double> &hv() const { return hv_; }

  private:
    void check(int rc, const char *what) const;
    void download();
    void run_sanity_checks(double time, std::size_t step);

    const SimulationConfig cfg_;
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

// solver_assembly.hpp:124-160: the fixed scheme of the optimized backends, and the config
// settings it ignores (printed to stderr).
void warn_if_ignored_optimized_settings(const SimulationConfig &cfg, const std::string &backend);

// time_utils.hpp:40-71
double estimate_eta_seconds(std::size_t completed, std::size_t total, double elapsed);
std::string format_duration(double seconds);
