// config.hpp -- simulation configuration of the HIP backend's host application.
//
// The schema is the reference's (configs/config.hpp:7-105, TOML keys of
// configs/simulation_config.toml:1-64) so that a reference configuration file runs unchanged
// here; the backend enum gains HIP (TOML value "HIP") for this backend.
#pragma once

#include <cstddef>
#include <filesystem>
#include <string>

// configs/config.hpp:7, plus HIP (the optimized scheme, the drop-in beside CUDA) and
// HIP_Serial (the modular Serial backend's scheme family on the GPU)
enum class BackendType { Serial, OptimizedSerial, OpenMP, CUDA, CUDA_4, HIP, HIP_Serial };

// configs/config.hpp:9, plus Transmissive (zero-gradient ghost copy; an extension honoured
// by the HIP backend's optimized scheme)
enum class BoundaryType { ReflectingWalls, Transmissive };

enum class BathymetryType { None, Flat, GaussHill };

enum class InitialConditionType { GaussInitial, StillWater, DamBreak, DamBreakRadial };

enum class ReconstructionType { PiecewiseConst, MUSCL };

enum class RiemannType { Rusanov, HLL, ROE };

enum class TimeIntegratorType { SSPRK3 };

struct MeshConfig {
    double Lx{};
    double Ly{};
    int nG{};
    int Nx{};
    int Ny{};
    std::string spatial_unit_x{};
    std::string spatial_unit_y{};
    std::string spatial_unit_h{};
};

struct BathymetryConfig {
    BathymetryType type{};
    double b0{};
    double bathy_x0{};
    double bathy_y0{};
    double bathy_sigma_x{};
    double bathy_sigma_y{};
    double bathy_peak_height{};
};

struct BoundaryConfig {
    BoundaryType type{};
};

struct InitialConditionConfig {
    InitialConditionType type{};
    double peak_height{};
    double x0{};
    double y0{};
    double sigma_x{};
    double sigma_y{};
    double h0{};
    double dam_height{};
    double dam_x{};
    double dam_x0{};
    double dam_y0{};
    double dam_radius{};
};

struct SolverConfig {
    ReconstructionType reconstruction{};
    RiemannType riemann{};
    TimeIntegratorType time{};
};

struct TimeConfig {
    double end_time{};
    std::size_t time_steps{};
    double cfl{};
    int save_every{};
    std::string time_unit{};
};

struct BackendConfig {
    BackendType type{};
    int threads{};
    int device{0};  // HIP device index ([backend] device, optional)
};

struct OutputConfig {
    std::filesystem::path path;
    bool compute_eta{};
};

struct SanityCheckConfig {
    bool mass_conservation{false};
    double mass_threshold{};
    bool positivity{false};
    bool debug{false};
    std::filesystem::path output_path;
};

struct SimulationConfig {
    MeshConfig mesh;
    BathymetryConfig bathymetry;
    BoundaryConfig boundary;
    InitialConditionConfig initial_condition;
    SolverConfig solver;
    TimeConfig time;
    BackendConfig backend;
    OutputConfig output;
    SanityCheckConfig sanity_checks;
};

// Reads a configuration file with the reference's keys (configs/simulation_config.cpp:75-200):
// every key the reference requires is required here, with the same optional keys and
// defaults.  Throws std::runtime_error naming the offending key or value.
SimulationConfig load_config(const std::filesystem::path &path);

// configs/parse_config_helper.hpp:5-31
std::string backend_name_from_cfg(const SimulationConfig &cfg);
std::string bathymetry_name_from_cfg(const SimulationConfig &cfg);
