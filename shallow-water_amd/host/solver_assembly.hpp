// solver_assembly.hpp -- backend dispatch of the host application (the reference's
// SolverAssembly, include/solver_assembly/solver_assembly.hpp:29-120): load and validate the
// configuration, then run the backend [backend] type names.  This build carries the HIP
// backend; the reference's other backends are reported as not built in.
#pragma once

#include <filesystem>

#include "config.hpp"

class SolverAssembly {
  public:
    explicit SolverAssembly(const std::filesystem::path &path);
    void run();
    const SimulationConfig &config() const { return cfg_; }

  private:
    SimulationConfig cfg_;
};
