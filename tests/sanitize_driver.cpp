// sanitize_driver.cpp -- drives the CPU oracle (oracle/swe_oracle.c, oracle/swe_modular.c)
// and the dual-number tangent (oracle/swe_tangent.cpp) over small grids of odd shapes,
// both boundary kinds and ghost widths 2 and 3, for a build with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_sanitizers.py).  TEST INFRASTRUCTURE ONLY.
// Exit status 0 when every call returned its expected status; the sanitizers report any
// out-of-bounds access, leak or undefined operation on stderr and abort.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

extern "C" {
struct oracle_grid {
    int32_t nx, ny, ng;
    double dx, dy, g;
    int32_t bc_kind;
};
int oracle_rhs(const oracle_grid *, double *, double *, double *, const double *, double *,
               double *, double *);
int oracle_run(const oracle_grid *, double *, double *, double *, const double *, int, double,
               long, double *, long *, int *);
double oracle_stable_dt(const oracle_grid *, const double *, const double *, const double *,
                        double, double *);
void oracle_apply_bc(const oracle_grid *, double *, double *, double *);
int oracle_mod_run(const oracle_grid *, int, int, int, double *, double *, double *,
                   const double *, int64_t, double);
void oracle_mod_apply_bc(const oracle_grid *, double *, double *, double *);
int tangent_run(int, int, int, double, double, double, int, double *, double *, double *,
                const double *, double *, double *, double *, const double *, const double *,
                long);
}

namespace {

uint64_t lcg = 0x9E3779B97F4A7C15ull;
double unif() {
    lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(lcg >> 11) * 0x1p-53;
}

struct Fields {
    std::vector<double> h, hu, hv, B;
};

// A wet state with bathymetry; wet_dry leaves some cells at depth 0.
Fields make_state(int nx, int ny, int ng, bool wet_dry) {
    const size_t n = (size_t)(nx + 2 * ng) * (ny + 2 * ng);
    Fields f{std::vector<double>(n), std::vector<double>(n), std::vector<double>(n),
             std::vector<double>(n)};
    for (size_t k = 0; k < n; ++k) {
        f.B[k] = 0.2 * unif();
        f.h[k] = wet_dry && unif() < 0.2 ? 0.0 : 0.5 + unif();
        f.hu[k] = 0.1 * (unif() - 0.5);
        f.hv[k] = 0.1 * (unif() - 0.5);
    }
    return f;
}

int failures = 0;
void expect(bool ok, const char *what, int nx, int ny, int ng, int bc) {
    if (!ok) {
        std::fprintf(stderr, "FAILED %s at %dx%d ng=%d bc=%d\n", what, nx, ny, ng, bc);
        ++failures;
    }
}

}  // namespace

int main() {
    const int shapes[][2] = {{3, 2}, {2, 7}, {17, 13}, {33, 64}, {61, 5}};
    for (const auto &sh : shapes)
        for (int ng = 2; ng <= 3; ++ng)
            for (int bc = 0; bc <= 1; ++bc)
                for (int wd = 0; wd <= 1; ++wd) {
                    const int nx = sh[0], ny = sh[1];
                    oracle_grid gr{nx, ny, ng, 1.0 / nx, 1.0 / ny, 9.80665, bc};
                    Fields f = make_state(nx, ny, ng, wd != 0);
                    oracle_apply_bc(&gr, f.B.data(), nullptr, nullptr);
                    const size_t n = f.h.size();
                    std::vector<double> r0(n), r1(n), r2(n);
                    expect(oracle_rhs(&gr, f.h.data(), f.hu.data(), f.hv.data(), f.B.data(),
                                      r0.data(), r1.data(), r2.data()) == 0,
                           "oracle_rhs", nx, ny, ng, bc);
                    double ms = 0.0;
                    const double dt = oracle_stable_dt(&gr, f.h.data(), f.hu.data(), f.hv.data(),
                                                       0.3, &ms);
                    expect(std::isfinite(dt) && dt > 0.0, "oracle_stable_dt", nx, ny, ng, bc);
                    std::vector<double> hist(6);
                    long done = 0;
                    int failed_stage = -1;
                    const int rc = oracle_run(&gr, f.h.data(), f.hu.data(), f.hv.data(),
                                              f.B.data(), 1, 0.3, 6, hist.data(), &done,
                                              &failed_stage);
                    expect((rc == 0 && failed_stage == 0) ||
                               (rc == 1 && failed_stage >= 1 && failed_stage <= 3),
                           "oracle_run", nx, ny, ng, bc);
                    if (ng == 2 && bc == 0) {
                        // the modular family: every reconstruction x Riemann combination
                        for (int bathy = 0; bathy <= 1; ++bathy)
                            for (int recon = 0; recon <= 1; ++recon)
                                for (int rie = 0; rie <= 2; ++rie) {
                                    Fields m = make_state(nx, ny, ng, false);
                                    const int rm = oracle_mod_run(&gr, bathy, recon, rie,
                                                                  m.h.data(), m.hu.data(),
                                                                  m.hv.data(), m.B.data(), 3,
                                                                  0.1 / (nx + ny));
                                    expect(rm >= 0 && rm <= 2, "oracle_mod_run", nx, ny, ng, bc);
                                }
                        // the tangent along a random direction
                        Fields t = make_state(nx, ny, ng, false);
                        std::vector<double> dh(n), dhu(n), dhv(n), dB(n), dts(4, 0.2 / (nx + ny));
                        for (size_t k = 0; k < n; ++k) {
                            dh[k] = unif() - 0.5;
                            dhu[k] = unif() - 0.5;
                            dhv[k] = unif() - 0.5;
                            dB[k] = unif() - 0.5;
                        }
                        const int rt = tangent_run(nx, ny, ng, gr.dx, gr.dy, gr.g, bc,
                                                   t.h.data(), t.hu.data(), t.hv.data(),
                                                   t.B.data(), dh.data(), dhu.data(), dhv.data(),
                                                   dB.data(), dts.data(), 4);
                        expect(rt == 0, "tangent_run", nx, ny, ng, bc);
                    }
                }
    std::printf("sanitize driver: %d failures\n", failures);
    return failures == 0 ? 0 : 1;
}
