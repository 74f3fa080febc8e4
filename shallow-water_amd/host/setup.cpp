// setup.cpp -- see setup.hpp.  The generator formulas are the reference's (cited per
// function) so that a configuration file produces the same initial arrays bit for bit.
#include "setup.hpp"

#include <cmath>
#include <stdexcept>
#include <string>

HostGrid make_grid(const SimulationConfig &cfg) {
    return HostGrid{cfg.mesh.Nx, cfg.mesh.Ny, cfg.mesh.nG, cfg.mesh.Lx, cfg.mesh.Ly};
}

void reflect_ghosts(const HostGrid &g, double *h, double *hu, double *hv) {
    const int nG = g.nG, Nx = g.Nx, Ny = g.Ny;
    for (int k = 0; k < nG; ++k) {
        const int iL = nG - 1 - k, sL = nG + k, iR = nG + Nx + k, sR = nG + Nx - 1 - k;
        for (int j = nG; j < nG + Ny; ++j) {
            h[g.idx(iL, j)] = h[g.idx(sL, j)];
            h[g.idx(iR, j)] = h[g.idx(sR, j)];
            if (hu) {
                hu[g.idx(iL, j)] = -hu[g.idx(sL, j)];
                hu[g.idx(iR, j)] = -hu[g.idx(sR, j)];
                hv[g.idx(iL, j)] = hv[g.idx(sL, j)];
                hv[g.idx(iR, j)] = hv[g.idx(sR, j)];
            }
        }
    }
    for (int k = 0; k < nG; ++k) {
        const int jB = nG - 1 - k, sB = nG + k, jT = nG + Ny + k, sT = nG + Ny - 1 - k;
        for (int i = 0; i < g.Nx_total(); ++i) {
            h[g.idx(i, jB)] = h[g.idx(i, sB)];
            h[g.idx(i, jT)] = h[g.idx(i, sT)];
            if (hu) {
                hu[g.idx(i, jB)] = hu[g.idx(i, sB)];
                hu[g.idx(i, jT)] = hu[g.idx(i, sT)];
                hv[g.idx(i, jB)] = -hv[g.idx(i, sB)];
                hv[g.idx(i, jT)] = -hv[g.idx(i, sT)];
            }
        }
    }
}

std::vector<double> make_bathymetry(const SimulationConfig &cfg, const HostGrid &g) {
    std::vector<double> B(g.n_total(), 0.0);
    const BathymetryConfig &b = cfg.bathymetry;
    switch (b.type) {
        case BathymetryType::Flat:  // flat_bathymetry.hpp: every cell, ghosts included
            for (double &v : B) v = b.b0;
            break;
        case BathymetryType::None:
            break;
        case BathymetryType::GaussHill:  // gaussian_hill.hpp:19-29: interior cells
            for (int i = 0; i < g.Nx; ++i)
                for (int j = 0; j < g.Ny; ++j) {
                    const double x = g.x_center(i), y = g.y_center(j);
                    B[g.idx(i + g.nG, j + g.nG)] =
                        b.b0 + b.bathy_peak_height *
                                   std::exp(-((x - b.bathy_x0) * (x - b.bathy_x0)) /
                                                (2.0 * b.bathy_sigma_x * b.bathy_sigma_x) -
                                            ((y - b.bathy_y0) * (y - b.bathy_y0)) /
                                                (2.0 * b.bathy_sigma_y * b.bathy_sigma_y));
                }
            break;
    }
    reflect_ghosts(g, B.data(), nullptr, nullptr);
    return B;
}

void make_initial_state(const SimulationConfig &cfg, const HostGrid &g, std::vector<double> &h,
                        std::vector<double> &hu, std::vector<double> &hv) {
    h.assign(g.n_total(), 0.0);
    hu.assign(g.n_total(), 0.0);
    hv.assign(g.n_total(), 0.0);
    const InitialConditionConfig &c = cfg.initial_condition;
    for (int i = 0; i < g.Nx; ++i)
        for (int j = 0; j < g.Ny; ++j) {
            const double x = g.x_center(i), y = g.y_center(j);
            double v = 0.0;
            switch (c.type) {
                case InitialConditionType::GaussInitial:  // gauss_initial.hpp:17-25
                    v = c.h0 + c.peak_height * std::exp(-((x - c.x0) * (x - c.x0)) /
                                                            (2 * c.sigma_x * c.sigma_x) -
                                                        ((y - c.y0) * (y - c.y0)) /
                                                            (2 * c.sigma_y * c.sigma_y));
                    break;
                case InitialConditionType::StillWater:  // still_water.hpp:10-16
                    v = c.h0;
                    break;
                case InitialConditionType::DamBreak:  // dam_break.hpp:14-26: h0 below dam_x
                    v = x < c.dam_x ? c.h0 : c.dam_height;
                    break;
                case InitialConditionType::DamBreakRadial: {  // dam_break_radial.hpp:16-28
                    const double dx = x - c.dam_x0, dy = y - c.dam_y0;
                    v = (dx * dx + dy * dy <= c.dam_radius * c.dam_radius) ? c.dam_height : c.h0;
                    break;
                }
            }
            h[g.idx(i + g.nG, j + g.nG)] = v;
        }
    reflect_ghosts(g, h.data(), hu.data(), hv.data());
}

void validate_config(const SimulationConfig &cfg) {
    auto need = [](bool ok, const char *msg) {
        if (!ok) throw std::runtime_error(msg);
    };
    const auto &m = cfg.mesh;
    const auto &t = cfg.time;
    const auto &ic = cfg.initial_condition;
    const auto &b = cfg.bathymetry;
    need(m.Nx > 0 && m.Ny > 0, "Nx and Ny must be positive");
    need(m.nG >= 1, "nG must be at least 1");
    need(t.end_time > 0.0, "end_time must be positive");
    need(t.time_steps > 0, "time_steps must be > 0");
    if (ic.type == InitialConditionType::GaussInitial || ic.type == InitialConditionType::StillWater)
        need(ic.h0 >= 0.0, "h0 must be positive");
    need(!(cfg.solver.reconstruction == ReconstructionType::MUSCL && m.nG < 2),
         "MUSCL reconstruction requires nG >= 2");
    need(m.Lx > 0.0 && m.Ly > 0.0, "Lx and Ly must be positive");
    need(t.cfl > 0.0, "cfl must be positive");
    need(t.cfl <= 1.0, "cfl should usually be <= 1.0");
    need(t.save_every > 0, "save_every must be > 0");
    need((std::size_t)t.save_every <= t.time_steps,
         "save_every must not be larger than time_steps");
    need(!(cfg.backend.type == BackendType::HIP && m.nG < 2), "HIP solver requires nG >= 2");
    if (b.type == BathymetryType::Flat || b.type == BathymetryType::GaussHill)
        need(std::isfinite(b.b0), "bathymetry.b0 must be finite");
    if (b.type == BathymetryType::GaussHill) {
        need(b.bathy_sigma_x > 0.0 && b.bathy_sigma_y > 0.0,
             "Gaussian bathymetry sigmas must be positive");
        need(b.bathy_peak_height >= 0.0, "bathymetry.bathy_peak_height must be non-negative");
        need(b.bathy_x0 >= 0.0 && b.bathy_x0 <= m.Lx && b.bathy_y0 >= 0.0 && b.bathy_y0 <= m.Ly,
             "Gaussian bathymetry center must lie inside the domain");
    }
    need(ic.h0 > 0.0, "initial_condition.h0 must be positive");
    if (ic.type == InitialConditionType::GaussInitial) {
        need(ic.sigma_x > 0.0 && ic.sigma_y > 0.0,
             "Gaussian initial-condition sigmas must be positive");
        need(ic.x0 >= 0.0 && ic.x0 <= m.Lx && ic.y0 >= 0.0 && ic.y0 <= m.Ly,
             "Gaussian initial-condition center must lie inside the domain");
        need(ic.h0 + ic.peak_height > 0.0,
             "Gaussian initial condition can produce non-positive water height: "
             "h0 + peak_height <= 0");
    }
    if (ic.type == InitialConditionType::DamBreak) {
        need(ic.dam_height > 0.0, "dam_height must be positive");
        need(ic.dam_x >= 0.0 && ic.dam_x <= m.Lx, "dam_x must lie inside the domain");
    }
    if (ic.type == InitialConditionType::DamBreakRadial) {
        need(ic.dam_height > 0.0, "dam_height must be positive");
        need(ic.dam_radius > 0.0, "dam_radius must be positive");
        need(ic.dam_x0 >= 0.0 && ic.dam_x0 <= m.Lx && ic.dam_y0 >= 0.0 && ic.dam_y0 <= m.Ly,
             "radial dam-break center must lie inside the domain");
    }
}
