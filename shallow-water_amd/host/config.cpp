// config.cpp -- load_config over toml_lite (reference behaviour:
// configs/simulation_config.cpp:11-200, configs/parse_config_helper.hpp:5-31).
#include "config.hpp"

#include <stdexcept>
#include <string>

#include "toml_lite.hpp"

namespace {

using toml_lite::Table;

template <class E>
struct Named {
    const char *name;
    E value;
};

template <class E, std::size_t N>
E parse_enum(const std::string &s, const Named<E> (&names)[N], const char *what) {
    for (const auto &n : names)
        if (s == n.name) return n.value;
    throw std::runtime_error(std::string("Unknown ") + what + ": " + s);
}

const Named<BackendType> kBackends[] = {
    {"Serial", BackendType::Serial}, {"OptimizedSerial", BackendType::OptimizedSerial},
    {"OpenMP", BackendType::OpenMP}, {"CUDA", BackendType::CUDA},
    {"CUDA_4", BackendType::CUDA_4}, {"HIP", BackendType::HIP},
    {"HIP_Serial", BackendType::HIP_Serial}};
const Named<BoundaryType> kBoundaries[] = {{"ReflectingWalls", BoundaryType::ReflectingWalls},
                                           {"Transmissive", BoundaryType::Transmissive}};
const Named<ReconstructionType> kRecon[] = {{"PiecewiseConst", ReconstructionType::PiecewiseConst},
                                            {"MUSCL", ReconstructionType::MUSCL}};
const Named<RiemannType> kRiemann[] = {
    {"Rusanov", RiemannType::Rusanov}, {"HLL", RiemannType::HLL}, {"ROE", RiemannType::ROE}};
const Named<TimeIntegratorType> kTime[] = {{"SSPRK3", TimeIntegratorType::SSPRK3}};
const Named<InitialConditionType> kIC[] = {
    {"GaussInitial", InitialConditionType::GaussInitial},
    {"StillWater", InitialConditionType::StillWater},
    {"DamBreak", InitialConditionType::DamBreak},
    {"DamBreakRadial", InitialConditionType::DamBreakRadial}};
const Named<BathymetryType> kBathy[] = {{"Flat", BathymetryType::Flat},
                                        {"None", BathymetryType::None},
                                        {"GaussHill", BathymetryType::GaussHill}};

[[noreturn]] void missing(const std::string &key) {
    throw std::runtime_error("Missing or invalid TOML value for key '" + key + "'");
}

const Table &table(const toml_lite::Document &doc, const std::string &name) {
    const auto it = doc.find(name);
    if (it == doc.end()) throw std::runtime_error("Missing or invalid [" + name + "] table");
    return it->second;
}

double req_double(const Table &t, const std::string &table_name, const std::string &key) {
    if (auto v = toml_lite::get_double(t, key)) return *v;
    missing(table_name + "." + key);
}

std::int64_t req_int(const Table &t, const std::string &table_name, const std::string &key) {
    if (auto v = toml_lite::get_int(t, key)) return *v;
    missing(table_name + "." + key);
}

std::string req_string(const Table &t, const std::string &table_name, const std::string &key) {
    if (auto v = toml_lite::get_string(t, key)) return *v;
    missing(table_name + "." + key);
}

}  // namespace

SimulationConfig load_config(const std::filesystem::path &path) {
    toml_lite::Document doc;
    try {
        doc = toml_lite::parse_file(path.string());
    } catch (const std::runtime_error &e) {
        throw std::runtime_error("Failed to parse TOML file '" + path.string() + "': " + e.what());
    }
    SimulationConfig cfg{};

    const Table &mesh = table(doc, "mesh");
    cfg.mesh.Lx = req_double(mesh, "mesh", "Lx");
    cfg.mesh.Ly = req_double(mesh, "mesh", "Ly");
    cfg.mesh.nG = static_cast<int>(req_int(mesh, "mesh", "nG"));
    cfg.mesh.Nx = static_cast<int>(req_int(mesh, "mesh", "Nx"));
    cfg.mesh.Ny = static_cast<int>(req_int(mesh, "mesh", "Ny"));
    cfg.mesh.spatial_unit_x = req_string(mesh, "mesh", "spatial_unit_x");
    cfg.mesh.spatial_unit_y = req_string(mesh, "mesh", "spatial_unit_y");
    cfg.mesh.spatial_unit_h = req_string(mesh, "mesh", "spatial_unit_h");

    const Table &bathy = table(doc, "bathymetry");
    cfg.bathymetry.type = parse_enum(req_string(bathy, "bathymetry", "type"), kBathy,
                                     "bathymetry type");
    cfg.bathymetry.b0 = req_double(bathy, "bathymetry", "b0");
    cfg.bathymetry.bathy_x0 = req_double(bathy, "bathymetry", "bathy_x0");
    cfg.bathymetry.bathy_y0 = req_double(bathy, "bathymetry", "bathy_y0");
    cfg.bathymetry.bathy_sigma_x = req_double(bathy, "bathymetry", "bathy_sigma_x");
    cfg.bathymetry.bathy_sigma_y = req_double(bathy, "bathymetry", "bathy_sigma_y");
    cfg.bathymetry.bathy_peak_height = req_double(bathy, "bathymetry", "bathy_peak_height");

    const Table &boundary = table(doc, "boundary");
    cfg.boundary.type = parse_enum(req_string(boundary, "boundary", "type"), kBoundaries,
                                   "boundary type");

    const Table &ic = table(doc, "initial_condition");
    const std::string icn = "initial_condition";
    cfg.initial_condition.type = parse_enum(req_string(ic, icn, "type"), kIC, "initial condition");
    cfg.initial_condition.peak_height = req_double(ic, icn, "peak_height");
    cfg.initial_condition.x0 = req_double(ic, icn, "x0");
    cfg.initial_condition.y0 = req_double(ic, icn, "y0");
    cfg.initial_condition.sigma_x = req_double(ic, icn, "sigma_x");
    cfg.initial_condition.sigma_y = req_double(ic, icn, "sigma_y");
    cfg.initial_condition.h0 = req_double(ic, icn, "h0");
    cfg.initial_condition.dam_height = req_double(ic, icn, "dam_height");
    cfg.initial_condition.dam_x = req_double(ic, icn, "dam_x");
    cfg.initial_condition.dam_x0 = req_double(ic, icn, "dam_x0");
    cfg.initial_condition.dam_y0 = req_double(ic, icn, "dam_y0");
    cfg.initial_condition.dam_radius = req_double(ic, icn, "dam_radius");

    const Table &solver = table(doc, "solver");
    cfg.solver.reconstruction = parse_enum(req_string(solver, "solver", "reconstruction"), kRecon,
                                           "reconstruction type");
    cfg.solver.riemann = parse_enum(req_string(solver, "solver", "riemann"), kRiemann,
                                    "riemann solver");
    cfg.solver.time = parse_enum(req_string(solver, "solver", "time"), kTime, "time integrator");

    const Table &time = table(doc, "time");
    cfg.time.end_time = req_double(time, "time", "end_time");
    const std::int64_t steps = req_int(time, "time", "time_steps");
    if (steps < 0) missing("time.time_steps");
    cfg.time.time_steps = static_cast<std::size_t>(steps);
    cfg.time.cfl = req_double(time, "time", "cfl");
    cfg.time.save_every = static_cast<int>(req_int(time, "time", "save_every"));
    cfg.time.time_unit = req_string(time, "time", "time_unit");

    const Table &backend = table(doc, "backend");
    cfg.backend.type = parse_enum(req_string(backend, "backend", "type"), kBackends,
                                  "backend type");
    cfg.backend.threads = static_cast<int>(req_int(backend, "backend", "threads"));
    if (auto d = toml_lite::get_int(backend, "device")) cfg.backend.device = static_cast<int>(*d);

    const Table &output = table(doc, "output");
    cfg.output.path = req_string(output, "output", "path");
    cfg.output.compute_eta = toml_lite::get_bool(output, "compute_eta").value_or(false);

    const auto sit = doc.find("sanity_checks");
    if (sit != doc.end()) {
        const Table &s = sit->second;
        cfg.sanity_checks.mass_conservation =
            toml_lite::get_bool(s, "mass_conservation").value_or(false);
        cfg.sanity_checks.mass_threshold = req_double(s, "sanity_checks", "mass_threshold");
        cfg.sanity_checks.positivity = toml_lite::get_bool(s, "positivity").value_or(false);
        cfg.sanity_checks.debug = toml_lite::get_bool(s, "debug").value_or(false);
        cfg.sanity_checks.output_path = req_string(s, "sanity_checks", "output_path");
    }
    return cfg;
}

std::string backend_name_from_cfg(const SimulationConfig &cfg) {
    switch (cfg.backend.type) {
        case BackendType::Serial: return "Serial";
        case BackendType::OptimizedSerial: return "OptimizedSerial";
        case BackendType::OpenMP: return "OpenMP";
        case BackendType::CUDA: return "CUDA";
        case BackendType::CUDA_4: return "CUDA_4";
        case BackendType::HIP: return "HIP";
        case BackendType::HIP_Serial: return "HIP_Serial";
    }
    return "UnknownBackend";
}

std::string bathymetry_name_from_cfg(const SimulationConfig &cfg) {
    switch (cfg.bathymetry.type) {
        case BathymetryType::None: return "None";
        case BathymetryType::Flat: return "Flat";
        case BathymetryType::GaussHill: return "Gaussian Hill";
    }
    return "UnknownBathymetry";
}
