// hip_solver.cpp -- see hip_solver.hpp.
#include "hip_solver.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <limits>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "swe_hip.h"

namespace {

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

constexpr double kGravity = 9.80665;  // include/constants.hpp:4

}  // namespace

void HIPSolverTimingStats::print_summary() const {
    const double total =
        host_to_device + device_to_host + steps + output + sanity_checks;
    std::cout << "\n========== HIP solver timing summary ==========\n";
    std::cout << "time steps           : " << time_steps << '\n';
    std::cout << "CFL check            : " << cfl_check << " s\n";
    std::cout << "measured solver time : " << total << " s\n";
    if (total <= 0.0) {
        std::cout << "No positive measured solver time.\n";
        std::cout << "===============================================\n";
        return;
    }
    const auto row = [total](const char *name, double s) {
        std::cout << std::left << std::setw(24) << name << std::right << std::setw(12) << s
                  << " s  " << std::setw(8) << 100.0 * s / total << " %\n";
    };
    std::cout << std::fixed << std::setprecision(6);
    row("host to device", host_to_device);
    row("device to host", device_to_host);
    row("fused RK3 steps", steps);
    row("output", output);
    row("sanity checks", sanity_checks);
    std::cout << "===============================================\n";
}

double estimate_eta_seconds(std::size_t completed, std::size_t total, double elapsed) {
    if (completed == 0 || completed >= total) return 0.0;
    return elapsed / (double)completed * (double)(total - completed);
}

std::string format_duration(double seconds) {
    if (seconds < 0.0) seconds = 0.0;
    const long long t = std::llround(seconds);
    const long long h = t / 3600, m = (t % 3600) / 60, s = t % 60;
    std::ostringstream o;
    if (h > 0)
        o << h << "h " << std::setw(2) << std::setfill('0') << m << "m " << std::setw(2) << s
          << "s";
    else if (m > 0)
        o << m << "m " << std::setw(2) << std::setfill('0') << s << "s";
    else
        o << s << "s";
    return o.str();
}

void warn_if_ignored_optimized_settings(const SimulationConfig &cfg, const std::string &backend) {
    bool warned = false;
    auto begin = [&]() {
        if (warned) return;
        std::cerr << "Warning: backend.type = \"" << backend
                  << "\" uses a fixed optimized solver configuration:\n"
                  << "  reconstruction = MUSCL\n"
                  << "  riemann        = HLL\n"
                  << "  time           = SSPRK3\n"
                  << "  boundary       = ReflectingWalls or Transmissive ([boundary] type)\n"
                  << "Some [solver] settings from the config may be ignored.\n\n";
        warned = true;
    };
    if (cfg.solver.reconstruction != ReconstructionType::MUSCL) {
        begin();
        std::cerr << "  Ignored reconstruction setting from config.\n";
    }
    if (cfg.solver.riemann != RiemannType::HLL) {
        begin();
        std::cerr << "  Ignored riemann setting from config.\n";
    }
    if (cfg.solver.time != TimeIntegratorType::SSPRK3) {
        begin();
        std::cerr << "  Ignored time-integrator setting from config.\n";
    }
    if (cfg.bathymetry.type == BathymetryType::None) {
        begin();
        std::cerr << "  Ignored bathymetry setting from config.\n"
                  << "For None-like Bathymetry choose flat with b0 = 0";
    }
}

static std::string boundary_name_of(const SimulationConfig &cfg) {
    return cfg.boundary.type == BoundaryType::Transmissive ? "Transmissive" : "Reflecting Walls";
}

static std::unique_ptr<SanityCheckNetCDFWriter> make_sanity_writer(const SimulationConfig &cfg,
                                                                   double dt,
                                                                   const std::string &riemann,
                                                                   const std::string &recon,
                                                                   const std::string &bathy) {
    if (!cfg.sanity_checks.debug) return nullptr;
    if (cfg.sanity_checks.output_path.empty())
        throw std::runtime_error("sanity_checks.output_path must be set if debug mode is enabled");
    return std::make_unique<SanityCheckNetCDFWriter>(
        cfg.sanity_checks.output_path.string(), cfg.time.time_unit, cfg.mesh.spatial_unit_h,
        cfg.time.save_every, dt, backend_name_from_cfg(cfg), riemann, recon, "SSPRK3",
        boundary_name_of(cfg), bathy);
}

// serial_solver.hpp:403-418
static std::string riemann_name_of(const HIPScheme &s) {
    if (s.scheme == SWE_SCHEME_OPTIMIZED || s.riemann == SWE_RIEMANN_HLL) return "HLL";
    return s.riemann == SWE_RIEMANN_RUSANOV ? "Rusanov" : "ROE";
}
static std::string recon_name_of(const HIPScheme &s) {
    if (s.scheme == SWE_SCHEME_OPTIMIZED || s.recon == SWE_RECON_MUSCL) return "MUSCL";
    return "PiecewiseConst";
}

static HIPScheme serial_scheme_of(const SimulationConfig &cfg) {
    HIPScheme s;
    s.scheme = cfg.bathymetry.type == BathymetryType::None ? SWE_SCHEME_SERIAL_PLAIN
                                                           : SWE_SCHEME_SERIAL;
    s.recon = cfg.solver.reconstruction == ReconstructionType::MUSCL ? SWE_RECON_MUSCL
                                                                     : SWE_RECON_PC;
    s.riemann = cfg.solver.riemann == RiemannType::HLL       ? SWE_RIEMANN_HLL
                : cfg.solver.riemann == RiemannType::Rusanov ? SWE_RIEMANN_RUSANOV
                                                             : SWE_RIEMANN_ROE;
    return s;
}

SerialHIPSolver::SerialHIPSolver(const SimulationConfig &cfg)
    : HIPSolverBase(cfg, serial_scheme_of(cfg)) {}

HIPSolverBase::HIPSolverBase(const SimulationConfig &cfg, HIPScheme scheme)
    : cfg_(cfg), scheme_(scheme), riemann_name_(riemann_name_of(scheme)),
      recon_name_(recon_name_of(scheme)), bathy_name_(bathymetry_name_from_cfg(cfg)),
      grid_(make_grid(cfg)),
      dt_(cfg.time.end_time / static_cast<double>(cfg.time.time_steps)),
      steps_(cfg.time.time_steps), save_every_(static_cast<std::size_t>(cfg.time.save_every)),
      writer_(cfg.output.path.string(), grid_, cfg.mesh.spatial_unit_x, cfg.mesh.spatial_unit_y,
              cfg.mesh.spatial_unit_h, cfg.time.time_unit, cfg.time.save_every),
      sanity_checks_(make_sanity_checks(cfg)),
      sanity_writer_(make_sanity_writer(cfg, dt_, riemann_name_, recon_name_, bathy_name_)) {
    if (grid_.nG < 2) throw std::runtime_error("The HIP solver requires nG >= 2 for MUSCL.");
    B_ = make_bathymetry(cfg_, grid_);
    writer_.write_bathymetry(B_.data());
    make_initial_state(cfg_, grid_, h_, hu_, hv_);
    for (auto &c : sanity_checks_) c->initialize(h_.data(), grid_);

    swe_grid_desc d{};
    d.nx = grid_.Nx;
    d.ny = grid_.Ny;
    d.ng = grid_.nG;
    d.dx = grid_.dx();
    d.dy = grid_.dy();
    d.g = kGravity;
    d.bc_kind = cfg_.boundary.type == BoundaryType::Transmissive ? SWE_BC_TRANSMISSIVE
                                                                 : SWE_BC_REFLECTING;
    check(swe_hip_create(&d, cfg_.backend.device, &ctx_), "swe_hip_create");
    if (scheme_.scheme != SWE_SCHEME_OPTIMIZED)
        check(swe_hip_set_scheme(ctx_, scheme_.scheme, scheme_.recon, scheme_.riemann),
              "swe_hip_set_scheme");
}

HIPSolverBase::~HIPSolverBase() { swe_hip_destroy(ctx_); }

// time_utils.hpp:13-36: min over interior cells of cfl / ((|u| + c)/dx + (|v| + c)/dy), with
// u = hu/h (no eps test); throws on h < 0.  The serial family's pre-loop CFL check.
double HIPSolverBase::serial_stable_dt() const {
    double lim = std::numeric_limits<double>::max();
    for (int i = grid_.nG; i < grid_.nG + grid_.Nx; ++i)
        for (int j = grid_.nG; j < grid_.nG + grid_.Ny; ++j) {
            const std::size_t k = grid_.idx(i, j);
            const double h = h_[k], hu = hu_[k], hv = hv_[k];
            if (h < 0.0)
                throw std::runtime_error("Could not perform stability check for dt because h <= 0!");
            const double u = hu / h, v = hv / h;
            const double c = cfg_.time.cfl / ((std::abs(u) + std::sqrt(kGravity * h)) / grid_.dx() +
                                              (std::abs(v) + std::sqrt(kGravity * h)) / grid_.dy());
            lim = std::min(lim, c);
        }
    return lim;
}

void HIPSolverBase::check(int rc, const char *what) const {
    if (rc == SWE_OK) return;
    std::string msg = std::string("HIP backend: ") + what + " failed: " + swe_hip_error_string(rc);
    if (ctx_) {
        const char *detail = swe_hip_last_error(ctx_);
        if (detail && *detail) msg += std::string(" (") + detail + ")";
    }
    throw std::runtime_error(msg);
}

void HIPSolverBase::download() {
    const auto t0 = Clock::now();
    check(swe_hip_download(ctx_, h_.data(), hu_.data(), hv_.data()), "swe_hip_download");
    timing_.device_to_host += since(t0);
}

void HIPSolverBase::run_sanity_checks(double time, std::size_t step) {
    const auto t0 = Clock::now();
    for (auto &c : sanity_checks_) c->evaluate(h_.data(), grid_, time, step, sanity_writer_.get());
    timing_.sanity_checks += since(t0);
}

void HIPSolverBase::run() {
    std::cout << std::unitbuf;
    std::cerr << std::unitbuf;
    const bool serial = scheme_.scheme != SWE_SCHEME_OPTIMIZED;
    if (serial)
        std::cout << "INFO: HIP serial-scheme solver active (" << recon_name_ << " + "
                  << riemann_name_ << ")\n";
    else
        std::cout << "INFO: HIP fused-stage solver active\n";

    {
        const auto t0 = Clock::now();
        check(swe_hip_upload(ctx_, h_.data(), hu_.data(), hv_.data(), B_.data()),
              "swe_hip_upload");
        timing_.host_to_device += since(t0);
    }
    {
        // cuda_solver.hpp:159-167: the fixed dt must satisfy the CFL limit of U^0
        const auto t0 = Clock::now();
        double dt_stable = 0.0;
        if (serial)
            dt_stable = serial_stable_dt();
        else
            check(swe_hip_stable_dt(ctx_, cfg_.time.cfl, &dt_stable, nullptr),
                  "swe_hip_stable_dt");
        timing_.cfl_check += since(t0);
        if (dt_ > dt_stable)
            throw std::runtime_error("Time step too large: dt = " + std::to_string(dt_) +
                                     ", but CFL requires dt <= " + std::to_string(dt_stable));
    }

    double time = 0.0;
    {
        const auto t0 = Clock::now();
        writer_.write_snapshot(h_.data(), hu_.data(), hv_.data(), time, dt_,
                               backend_name_from_cfg(cfg_), riemann_name_, recon_name_, "SSPRK3",
                               boundary_name_of(cfg_), bathy_name_);
        timing_.output += since(t0);
    }
    run_sanity_checks(time, 0);

    // serial_solver.hpp:181
    if (serial && scheme_.scheme == SWE_SCHEME_SERIAL) printf("\nINFO: Bathymetry enabled.\n\n");

    const bool estimate_eta = cfg_.output.compute_eta;
    const std::size_t eta_probe = std::min<std::size_t>(100, steps_);
    bool eta_printed = false;
    const auto wall0 = Clock::now();

    std::size_t done = 0;
    while (done < steps_) {
        // advance to the next save point (or the ETA probe) in one replayed-graph call
        std::size_t next = std::min(steps_, (done / save_every_ + 1) * save_every_);
        if (estimate_eta && !eta_printed && eta_probe > done && eta_probe < next) next = eta_probe;
        const std::size_t n = next - done;
        {
            const auto t0 = Clock::now();
            check(swe_hip_step(ctx_, dt_, (int64_t)n), "swe_hip_step");
            timing_.steps += since(t0);
        }
        // time accumulates step by step, as in the reference's loop
        for (std::size_t k = 0; k < n; ++k) time += dt_;
        done = next;
        timing_.time_steps = done;

        if (estimate_eta && !eta_printed && done == eta_probe) {
            std::cout << "ETA = "
                      << format_duration(estimate_eta_seconds(done, steps_, since(wall0)))
                      << '\n';
            eta_printed = true;
        }
        if (done % save_every_ == 0 || done == steps_) {
            download();
            const auto t0 = Clock::now();
            writer_.write_snapshot(h_.data(), hu_.data(), hv_.data(), time);
            timing_.output += since(t0);
            run_sanity_checks(time, done);
        }
    }
    timing_.print_summary();
}
