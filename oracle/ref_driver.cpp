// ref_driver.cpp -- C-ABI harness around the reference's own header-only kernels.
//
// TEST INFRASTRUCTURE ONLY (the checker's checker).  Built by oracle/Makefile target
// `ref` into oracle/_ref/libswe_ref.so, compiling the unmodified reference headers where
// they lie (-I<reference root>):
//   include/backend/openmp/omp_kernels.hpp       (namespace fast_hll_muscl_bathy_omp)
//   include/backend/serial/HPC/HPC_kernels.hpp   (namespace fast_hll_muscl_bathy)
// Nothing from the reference is copied here.  What this file adds is only the glue the
// reference keeps in classes that cannot be compiled without NetCDF
// (include/io/netCDF_writer.hpp:8): the rk3_step sequence of omp_solver.hpp:250-338 and
// the private ghost fill of omp_solver.hpp:340-432, restated so the reference kernels
// can be driven over raw arrays.  g is the reference's compile-time constant 9.80665.
// ref_initial_state runs the reference's own setup code (apply_bathymetry.hpp,
// apply_initial_condition.hpp) the way the solver constructor does
// (omp_solver.hpp:118-127), for the host application's initial-state parity.
#include "include/backend/openmp/omp_kernels.hpp"
#include "include/backend/serial/HPC/HPC_kernels.hpp"
#include "include/bathymetry/apply_bathymetry.hpp"
#include "include/initial_condition/apply_initial_condition.hpp"

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

namespace K = fast_hll_muscl_bathy_omp;
namespace S = fast_hll_muscl_bathy;

K::GridView make_view(int nx, int ny, int ng, double dx, double dy) {
    const int nxt = nx + 2 * ng, nyt = ny + 2 * ng;
    return K::GridView{nx, ny, ng, nxt, nyt, nyt, dx, dy};
}

S::GridView make_sview(int nx, int ny, int ng, double dx, double dy) {
    const int nxt = nx + 2 * ng, nyt = ny + 2 * ng;
    return S::GridView{nx, ny, ng, nxt, nyt, nyt, dx, dy};
}

// Restatement of the private reflecting fill, omp_solver.hpp:340-393 (a == nullptr for
// hu/hv gives the scalar fill of :395-432).
void reflect(const K::GridView &g, double *h, double *hu, double *hv) {
    const int nG = g.nG, Nx = g.Nx, Ny = g.Ny, s = g.stride;
    for (int gg = 0; gg < nG; ++gg) {
        const int iL = nG - 1 - gg, iSL = nG + gg, iR = nG + Nx + gg, iSR = nG + Nx - 1 - gg;
        for (int j = nG; j < nG + Ny; ++j) {
            const int kL = K::idx(iL, j, s), kSL = K::idx(iSL, j, s);
            const int kR = K::idx(iR, j, s), kSR = K::idx(iSR, j, s);
            h[kL] = h[kSL];
            h[kR] = h[kSR];
            if (hu) {
                hu[kL] = -hu[kSL];
                hv[kL] = hv[kSL];
                hu[kR] = -hu[kSR];
                hv[kR] = hv[kSR];
            }
        }
    }
    for (int gg = 0; gg < nG; ++gg) {
        const int jB = nG - 1 - gg, jSB = nG + gg, jT = nG + Ny + gg, jST = nG + Ny - 1 - gg;
        for (int i = 0; i < g.Nx_total; ++i) {
            const int kB = K::idx(i, jB, s), kSB = K::idx(i, jSB, s);
            const int kT = K::idx(i, jT, s), kST = K::idx(i, jST, s);
            h[kB] = h[kSB];
            h[kT] = h[kST];
            if (hu) {
                hu[kB] = hu[kSB];
                hv[kB] = -hv[kSB];
                hu[kT] = hu[kST];
                hv[kT] = -hv[kST];
            }
        }
    }
}

struct Work {
    std::size_t n;
    std::vector<double> h1, hu1, hv1, h2, hu2, hv2, rh, rhu, rhv;
    std::vector<double> f[12];
    explicit Work(std::size_t n_)
        : n(n_), h1(n_), hu1(n_), hv1(n_), h2(n_), hu2(n_), hv2(n_), rh(n_), rhu(n_), rhv(n_) {
        for (auto &a : f) a.assign(n_, 0.0);
    }
};

void rhs(const K::GridView &g, Work &w, const double *h, const double *hu, const double *hv,
         const double *B) {
    K::compute_x_fluxes(g, h, hu, hv, B, w.f[0].data(), w.f[1].data(), w.f[2].data(),
                        w.f[3].data(), w.f[4].data(), w.f[5].data());
    K::compute_y_fluxes(g, h, hu, hv, B, w.f[6].data(), w.f[7].data(), w.f[8].data(),
                        w.f[9].data(), w.f[10].data(), w.f[11].data());
    K::apply_divergence(g, w.f[0].data(), w.f[1].data(), w.f[2].data(), w.f[3].data(),
                        w.f[4].data(), w.f[5].data(), w.f[6].data(), w.f[7].data(),
                        w.f[8].data(), w.f[9].data(), w.f[10].data(), w.f[11].data(), w.rh.data(),
                        w.rhu.data(), w.rhv.data());
}

// omp_solver.hpp:250-338, restated over the reference kernels.
void rk3(const K::GridView &g, Work &w, double *h, double *hu, double *hv, const double *B,
         double dt) {
    const std::ptrdiff_t n = static_cast<std::ptrdiff_t>(w.n);
    reflect(g, h, hu, hv);
    rhs(g, w, h, hu, hv, B);
#pragma omp parallel for schedule(static)
    for (std::ptrdiff_t k = 0; k < n; ++k) {
        w.h1[k] = h[k] + dt * w.rh[k];
        w.hu1[k] = hu[k] + dt * w.rhu[k];
        w.hv1[k] = hv[k] + dt * w.rhv[k];
    }
    K::enforce_positivity(g, w.h1.data(), w.hu1.data(), w.hv1.data());
    reflect(g, w.h1.data(), w.hu1.data(), w.hv1.data());
    rhs(g, w, w.h1.data(), w.hu1.data(), w.hv1.data(), B);
#pragma omp parallel for schedule(static)
    for (std::ptrdiff_t k = 0; k < n; ++k) {
        w.h2[k] = 0.75 * h[k] + 0.25 * (w.h1[k] + dt * w.rh[k]);
        w.hu2[k] = 0.75 * hu[k] + 0.25 * (w.hu1[k] + dt * w.rhu[k]);
        w.hv2[k] = 0.75 * hv[k] + 0.25 * (w.hv1[k] + dt * w.rhv[k]);
    }
    K::enforce_positivity(g, w.h2.data(), w.hu2.data(), w.hv2.data());
    reflect(g, w.h2.data(), w.hu2.data(), w.hv2.data());
    rhs(g, w, w.h2.data(), w.hu2.data(), w.hv2.data(), B);
#pragma omp parallel for schedule(static)
    for (std::ptrdiff_t k = 0; k < n; ++k) {
        h[k] = (1.0 / 3.0) * h[k] + (2.0 / 3.0) * (w.h2[k] + dt * w.rh[k]);
        hu[k] = (1.0 / 3.0) * hu[k] + (2.0 / 3.0) * (w.hu2[k] + dt * w.rhu[k]);
        hv[k] = (1.0 / 3.0) * hv[k] + (2.0 / 3.0) * (w.hv2[k] + dt * w.rhv[k]);
    }
    K::enforce_positivity(g, h, hu, hv);
    reflect(g, h, hu, hv);
}

} // namespace

extern "C" {

double ref_g(void) { return constants::g; }

int ref_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

double ref_minmod(double a, double b) { return K::minmod(a, b); }
double hpc_minmod(double a, double b) { return S::minmod(a, b); }

void ref_hll_x(double hL, double huL, double hvL, double hR, double huR, double hvR,
               double *out) {
    K::hll_x(hL, huL, hvL, hR, huR, hvR, out[0], out[1], out[2]);
}
void ref_hll_y(double hB, double huB, double hvB, double hT, double huT, double hvT,
               double *out) {
    K::hll_y(hB, huB, hvB, hT, huT, hvT, out[0], out[1], out[2]);
}

// f: 12 arrays fxm(h,hu,hv) fxp(h,hu,hv) fym(..) fyp(..); rhs: 3 arrays.  All n_total.
void ref_faces_and_rhs(int nx, int ny, int ng, double dx, double dy, const double *h,
                       const double *hu, const double *hv, const double *B, double *const *f,
                       double *const *r) {
    const auto g = make_view(nx, ny, ng, dx, dy);
    K::compute_x_fluxes(g, h, hu, hv, B, f[0], f[1], f[2], f[3], f[4], f[5]);
    K::compute_y_fluxes(g, h, hu, hv, B, f[6], f[7], f[8], f[9], f[10], f[11]);
    K::apply_divergence(g, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], f[9], f[10],
                        f[11], r[0], r[1], r[2]);
}

// Same through the serial HPC twin (include/backend/serial/HPC/HPC_kernels.hpp).
void hpc_faces_and_rhs(int nx, int ny, int ng, double dx, double dy, const double *h,
                       const double *hu, const double *hv, const double *B, double *const *f,
                       double *const *r) {
    const auto g = make_sview(nx, ny, ng, dx, dy);
    S::compute_x_fluxes(g, h, hu, hv, B, f[0], f[1], f[2], f[3], f[4], f[5]);
    S::compute_y_fluxes(g, h, hu, hv, B, f[6], f[7], f[8], f[9], f[10], f[11]);
    S::apply_divergence(g, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], f[9], f[10],
                        f[11], r[0], r[1], r[2]);
}

double ref_stable_dt(int nx, int ny, int ng, double dx, double dy, const double *h,
                     const double *hu, const double *hv, double cfl) {
    return K::compute_stable_dt(make_view(nx, ny, ng, dx, dy), h, hu, hv, cfl);
}

// Returns 1 when the reference throws on a non-finite state.
int ref_enforce_positivity(int nx, int ny, int ng, double dx, double dy, double *h, double *hu,
                           double *hv) {
    try {
        K::enforce_positivity(make_view(nx, ny, ng, dx, dy), h, hu, hv);
    } catch (const std::runtime_error &) {
        return 1;
    }
    return 0;
}

void ref_reflect(int nx, int ny, int ng, double dx, double dy, double *h, double *hu,
                 double *hv) {
    reflect(make_view(nx, ny, ng, dx, dy), h, hu, hv);
}

// Advance nsteps fixed-dt (dt_mode 0) or per-step CFL (dt_mode 1, dt_or_cfl = cfl) steps.
// Returns 0, 1 on a non-finite state, -1 on any other exception.
int ref_run(int nx, int ny, int ng, double dx, double dy, double *h, double *hu, double *hv,
            const double *B, int dt_mode, double dt_or_cfl, long nsteps, double *dt_hist) {
    const auto g = make_view(nx, ny, ng, dx, dy);
    try {
        Work w(static_cast<std::size_t>(g.Nx_total) * static_cast<std::size_t>(g.Ny_total));
        for (long s = 0; s < nsteps; ++s) {
            double dt = dt_or_cfl;
            if (dt_mode == 1) dt = K::compute_stable_dt(g, h, hu, hv, dt_or_cfl);
            if (dt_hist) dt_hist[s] = dt;
            rk3(g, w, h, hu, hv, B, dt);
        }
    } catch (const std::runtime_error &) {
        return 1;
    } catch (...) {
        return -1;
    }
    return 0;
}

// The solver constructor's initial arrays (omp_solver.hpp:118-127): the reference's
// apply_bathymetry then the scalar ghost fill of B, apply_initial_condition then the ghost
// fill of (h, hu, hv).  ic = {peak_height, x0, y0, sigma_x, sigma_y, h0, dam_height, dam_x,
// dam_x0, dam_y0, dam_radius}; bt = {b0, bathy_x0, bathy_y0, bathy_sigma_x, bathy_sigma_y,
// bathy_peak_height}; the types are the reference enums' values.  Returns 0, or -1 when the
// reference throws (unsupported type).
int ref_initial_state(int nx, int ny, int ng, double lx, double ly, int ic_type,
                      const double *ic, int bathy_type, const double *bt, double *h,
                      double *hu, double *hv, double *B) {
    try {
        SimulationConfig cfg{};
        cfg.mesh.Lx = lx;
        cfg.mesh.Ly = ly;
        cfg.mesh.nG = ng;
        cfg.mesh.Nx = nx;
        cfg.mesh.Ny = ny;
        auto &c = cfg.initial_condition;
        c.type = static_cast<InitialConditionType>(ic_type);
        c.peak_height = ic[0];
        c.x0 = ic[1];
        c.y0 = ic[2];
        c.sigma_x = ic[3];
        c.sigma_y = ic[4];
        c.h0 = ic[5];
        c.dam_height = ic[6];
        c.dam_x = ic[7];
        c.dam_x0 = ic[8];
        c.dam_y0 = ic[9];
        c.dam_radius = ic[10];
        auto &b = cfg.bathymetry;
        b.type = static_cast<BathymetryType>(bathy_type);
        b.b0 = bt[0];
        b.bathy_x0 = bt[1];
        b.bathy_y0 = bt[2];
        b.bathy_sigma_x = bt[3];
        b.bathy_sigma_y = bt[4];
        b.bathy_peak_height = bt[5];
        const Grid grid(nx, ny, lx, ly, ng);
        const auto v = make_view(nx, ny, ng, lx / nx, ly / ny);
        const std::size_t n = static_cast<std::size_t>(grid.Nx_total()) * grid.Ny_total();
        Array2D Bt(grid.Nx_total(), grid.Ny_total());
        apply_bathymetry(cfg, grid, Bt);
        std::memcpy(B, Bt.data(), n * sizeof(double));
        reflect(v, B, nullptr, nullptr);
        State U(grid);
        apply_initial_condition(cfg, grid, U);
        std::memcpy(h, U.h().data(), n * sizeof(double));
        std::memcpy(hu, U.hu().data(), n * sizeof(double));
        std::memcpy(hv, U.hv().data(), n * sizeof(double));
        reflect(v, h, hu, hv);
    } catch (...) {
        return -1;
    }
    return 0;
}

} // extern "C"
