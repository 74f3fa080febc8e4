// sanity.hpp -- the reference's run-time sanity checks on downloaded states
// (tests/sanity_checks/{mass_conservation,positivity,debug}.hpp, factory in
// tests/sanity_checks/sanity_checks.cpp:10-25): relative mass drift against the initial mass
// (throws above the threshold), negative depths (throws), and the debug record of mass drift
// and minimum depth per saved step.
#pragma once

#include <memory>
#include <vector>

#include "netcdf_classic.hpp"

class SanityCheck {
  public:
    virtual ~SanityCheck() = default;
    virtual const char *name() const = 0;
    virtual void initialize(const double *h, const HostGrid &g) = 0;
    virtual void evaluate(const double *h, const HostGrid &g, double time, std::size_t step,
                          SanityCheckNetCDFWriter *writer) = 0;
};

std::vector<std::unique_ptr<SanityCheck>> make_sanity_checks(const SimulationConfig &cfg);

// sum over the interior of h * dx * dy, and the minimum interior h
double interior_mass(const double *h, const HostGrid &g, double *h_min = nullptr);
