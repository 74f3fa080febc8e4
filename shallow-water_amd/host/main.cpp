// main.cpp -- host application of the HIP backend (the reference's apps/main.cpp, which runs
// SolverAssembly on configs/simulation_config.toml).
//
//   swe_hip_app [CONFIG]                      run the configured simulation
//   swe_hip_app --print-config [CONFIG]       parse + validate, print the values (no GPU)
//   swe_hip_app --initial-state OUT [CONFIG]  write B and the initial snapshot to OUT (no GPU)
//
// CONFIG defaults to configs/simulation_config.toml.  Errors print "Error: ..." and exit 1.
#include <cstdlib>
#include <iostream>
#include <string>

#include "hip_solver.hpp"
#include "netcdf_classic.hpp"
#include "setup.hpp"
#include "solver_assembly.hpp"

namespace {

const char *kDefaultConfig = "configs/simulation_config.toml";

void print_config(const SimulationConfig &c) {
    std::cout.precision(17);
    std::cout << "mesh.Lx = " << c.mesh.Lx << "\nmesh.Ly = " << c.mesh.Ly
              << "\nmesh.nG = " << c.mesh.nG << "\nmesh.Nx = " << c.mesh.Nx
              << "\nmesh.Ny = " << c.mesh.Ny << "\nbathymetry.type = "
              << bathymetry_name_from_cfg(c) << "\nbathymetry.b0 = " << c.bathymetry.b0
              << "\ninitial_condition.type = " << static_cast<int>(c.initial_condition.type)
              << "\ninitial_condition.h0 = " << c.initial_condition.h0
              << "\ntime.end_time = " << c.time.end_time
              << "\ntime.time_steps = " << c.time.time_steps << "\ntime.cfl = " << c.time.cfl
              << "\ntime.save_every = " << c.time.save_every
              << "\nbackend.type = " << backend_name_from_cfg(c)
              << "\nbackend.threads = " << c.backend.threads
              << "\nbackend.device = " << c.backend.device
              << "\noutput.path = " << c.output.path.string()
              << "\noutput.compute_eta = " << (c.output.compute_eta ? "true" : "false")
              << "\nsanity_checks.mass_conservation = "
              << (c.sanity_checks.mass_conservation ? "true" : "false")
              << "\nsanity_checks.mass_threshold = " << c.sanity_checks.mass_threshold
              << "\nsanity_checks.positivity = " << (c.sanity_checks.positivity ? "true" : "false")
              << "\nsanity_checks.debug = " << (c.sanity_checks.debug ? "true" : "false") << '\n';
}

}  // namespace

int main(int argc, char **argv) {
    try {
        std::string mode, out, path = kDefaultConfig;
        int a = 1;
        if (a < argc && std::string(argv[a]) == "--print-config") {
            mode = "print";
            ++a;
        } else if (a < argc && std::string(argv[a]) == "--initial-state") {
            mode = "initial";
            if (++a >= argc) throw std::runtime_error("--initial-state needs an output path");
            out = argv[a++];
        }
        if (a < argc) path = argv[a++];
        if (a < argc) throw std::runtime_error(std::string("unexpected argument: ") + argv[a]);

        SolverAssembly app(path);
        if (mode == "print") {
            validate_config(app.config());
            print_config(app.config());
            return EXIT_SUCCESS;
        }
        if (mode == "initial") {
            const SimulationConfig &c = app.config();
            validate_config(c);
            const HostGrid g = make_grid(c);
            const std::vector<double> B = make_bathymetry(c, g);
            std::vector<double> h, hu, hv;
            make_initial_state(c, g, h, hu, hv);
            NetCDFWriter w(out, g, c.mesh.spatial_unit_x, c.mesh.spatial_unit_y,
                           c.mesh.spatial_unit_h, c.time.time_unit, c.time.save_every);
            w.write_bathymetry(B.data());
            w.write_snapshot(h.data(), hu.data(), hv.data(), 0.0,
                             c.time.end_time / static_cast<double>(c.time.time_steps),
                             backend_name_from_cfg(c), "HLL", "MUSCL", "SSPRK3",
                             "Reflecting Walls", bathymetry_name_from_cfg(c));
            return EXIT_SUCCESS;
        }
        app.run();
    } catch (const std::exception &e) {
        std::cerr << "Error: " << e.what() << '\n';
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
