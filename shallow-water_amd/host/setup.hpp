// setup.hpp -- grid geometry, bottom and initial state on the host, in the reference layout
// (row-major, k = i*(Ny+2nG) + j, i = x).
#pragma once

#include <vector>

#include "config.hpp"

// include/core/grid.hpp:3-31
struct HostGrid {
    int Nx, Ny, nG;
    double Lx, Ly;
    int Nx_total() const { return Nx + 2 * nG; }
    int Ny_total() const { return Ny + 2 * nG; }
    double dx() const { return Lx / Nx; }
    double dy() const { return Ly / Ny; }
    double x_center(int i) const { return (i + 0.5) * dx(); }
    double y_center(int j) const { return (j + 0.5) * dy(); }
    std::size_t n_total() const { return (std::size_t)Nx_total() * (std::size_t)Ny_total(); }
    std::size_t idx(int i, int j) const { return (std::size_t)i * Ny_total() + j; }
};

HostGrid make_grid(const SimulationConfig &cfg);

// include/bathymetry/apply_bathymetry.hpp:12-38 (Flat, GaussHill, None), then the scalar
// reflecting ghost fill the solvers apply to B (omp_solver.hpp:395-432).
std::vector<double> make_bathymetry(const SimulationConfig &cfg, const HostGrid &g);

// include/initial_condition/apply_initial_condition.hpp:13-43 (GaussInitial, StillWater,
// DamBreak, DamBreakRadial), then the reflecting ghost fill (omp_solver.hpp:340-393).
void make_initial_state(const SimulationConfig &cfg, const HostGrid &g, std::vector<double> &h,
                        std::vector<double> &hu, std::vector<double> &hv);

// Reflecting ghost fill: x walls over interior j (hu negated), then y walls over every i (hv
// negated); a null hu fills one scalar field without sign changes.
void reflect_ghosts(const HostGrid &g, double *h, double *hu, double *hv);

// The reference's configuration checks (solver_assembly.hpp:167-299).
void validate_config(const SimulationConfig &cfg);
