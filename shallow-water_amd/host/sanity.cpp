// sanity.cpp -- see sanity.hpp.
#include "sanity.hpp"

#include <cmath>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>

double interior_mass(const double *h, const HostGrid &g, double *h_min) {
    const double dxdy = g.dx() * g.dy();
    double mass = 0.0, mn = std::numeric_limits<double>::max();
    for (int i = g.nG; i < g.nG + g.Nx; ++i)
        for (int j = g.nG; j < g.nG + g.Ny; ++j) {
            const double v = h[g.idx(i, j)];
            mass += v * dxdy;
            if (v < mn) mn = v;
        }
    if (h_min) *h_min = mn;
    return mass;
}

namespace {

double rel_drift(double mass, double initial) {
    const double err = std::abs(mass - initial);
    return std::abs(initial) > 0.0 ? err / std::abs(initial) : err;
}

class MassConservation : public SanityCheck {
  public:
    explicit MassConservation(double threshold) : threshold_(threshold) {}
    const char *name() const override { return "MassConservation"; }
    void initialize(const double *h, const HostGrid &g) override {
        initial_ = interior_mass(h, g);
        ready_ = true;
        std::cout << "INFO: Sanity check for mass conservation active (threshold = " << threshold_
                  << ")\n";
    }
    void evaluate(const double *h, const HostGrid &g, double time, std::size_t step,
                  SanityCheckNetCDFWriter *) override {
        if (!ready_) initialize(h, g);
        const double e = rel_drift(interior_mass(h, g), initial_);
        if (e > threshold_)
            throw std::runtime_error("Mass conservation check failed: diverging rel mass err = " +
                                     std::to_string(e) + " at step = " + std::to_string(step) +
                                     ", time = " + std::to_string(time));
    }

  private:
    double threshold_, initial_ = 0.0;
    bool ready_ = false;
};

class Positivity : public SanityCheck {
  public:
    const char *name() const override { return "Positivity"; }
    void initialize(const double *h, const HostGrid &g) override {
        check(h, g, 0.0, 0);
        std::cout << "INFO: Sanity check for positivity active\n";
    }
    void evaluate(const double *h, const HostGrid &g, double time, std::size_t step,
                  SanityCheckNetCDFWriter *) override {
        check(h, g, time, step);
    }

  private:
    static void check(const double *h, const HostGrid &g, double time, std::size_t step) {
        for (int i = g.nG; i < g.nG + g.Nx; ++i)
            for (int j = g.nG; j < g.nG + g.Ny; ++j) {
                const double v = h[g.idx(i, j)];
                if (v < 0.0)
                    throw std::runtime_error(
                        "Positivity check failed: negative water height h = " + std::to_string(v) +
                        " at cell (" + std::to_string(i - g.nG) + ", " + std::to_string(j - g.nG) +
                        "), step = " + std::to_string(step) + ", time = " + std::to_string(time));
            }
    }
};

class Debug : public SanityCheck {
  public:
    const char *name() const override { return "Debug"; }
    void initialize(const double *h, const HostGrid &g) override {
        initial_ = interior_mass(h, g);
        ready_ = true;
    }
    void evaluate(const double *h, const HostGrid &g, double time, std::size_t step,
                  SanityCheckNetCDFWriter *writer) override {
        if (!ready_) initialize(h, g);
        if (writer == nullptr) throw std::runtime_error("Debug sanity check requires a sanity writer");
        double h_min = 0.0;
        const double mass = interior_mass(h, g, &h_min);
        writer->write(step, time, rel_drift(mass, initial_), h_min);
    }

  private:
    double initial_ = 0.0;
    bool ready_ = false;
};

}  // namespace

std::vector<std::unique_ptr<SanityCheck>> make_sanity_checks(const SimulationConfig &cfg) {
    std::vector<std::unique_ptr<SanityCheck>> checks;
    if (cfg.sanity_checks.mass_conservation)
        checks.push_back(std::make_unique<MassConservation>(cfg.sanity_checks.mass_threshold));
    if (cfg.sanity_checks.positivity) checks.push_back(std::make_unique<Positivity>());
    if (cfg.sanity_checks.debug) checks.push_back(std::make_unique<Debug>());
    return checks;
}
