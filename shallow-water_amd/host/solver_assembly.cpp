// solver_assembly.cpp -- see solver_assembly.hpp.
#include "solver_assembly.hpp"

#include <stdexcept>
#include <string>

#include "hip_solver.hpp"
#include "setup.hpp"

SolverAssembly::SolverAssembly(const std::filesystem::path &path) : cfg_(load_config(path)) {}

void SolverAssembly::run() {
    validate_config(cfg_);
    switch (cfg_.backend.type) {
        case BackendType::HIP: {
            warn_if_ignored_optimized_settings(cfg_, "HIP");
            if (cfg_.bathymetry.type == BathymetryType::Flat ||
                cfg_.bathymetry.type == BathymetryType::GaussHill) {
                FastHLLMUSCLBathyHIPSolver solver(cfg_);
                solver.run();
                return;
            }
            throw std::runtime_error(
                "Unsupported HIP solver combination. The HIP backend supports MUSCL + HLL + "
                "SSPRK3 + ReflectingWalls or Transmissive with Flat or GaussHill bathymetry.");
        }
        case BackendType::HIP_Serial: {
            // the combinations of the reference's run_serial (solver_assembly.hpp:315-500)
            const SolverConfig &sc = cfg_.solver;
            const bool walls = cfg_.boundary.type == BoundaryType::ReflectingWalls &&
                               sc.time == TimeIntegratorType::SSPRK3;
            const bool plain = cfg_.bathymetry.type == BathymetryType::None;
            const bool bathy = (cfg_.bathymetry.type == BathymetryType::Flat ||
                                cfg_.bathymetry.type == BathymetryType::GaussHill) &&
                               sc.riemann == RiemannType::HLL;
            if (walls && (plain || bathy)) {
                SerialHIPSolver solver(cfg_);
                solver.run();
                return;
            }
            throw std::runtime_error("Unsupported serial solver combination");
        }
        case BackendType::Serial:
        case BackendType::OptimizedSerial:
        case BackendType::OpenMP:
        case BackendType::CUDA:
        case BackendType::CUDA_4:
            throw std::runtime_error("backend type \"" + backend_name_from_cfg(cfg_) +
                                     "\" was requested, but this binary carries the HIP backend "
                                     "only. Set [backend] type = \"HIP\" or \"HIP_Serial\".");
    }
    throw std::runtime_error("Unsupported backend");
}
