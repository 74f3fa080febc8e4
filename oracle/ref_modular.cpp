// ref_modular.cpp -- C-ABI harness around the reference's own modular serial solver parts.
//
// TEST INFRASTRUCTURE ONLY (the checker of oracle/swe_modular.c).  Built by oracle/Makefile
// target `ref` into oracle/_ref/libswe_ref_modular.so from the unmodified reference headers
// where they lie (-I<reference root>):
//   include/core/{grid,state,cell_state,array2D,xflux_field,yflux_field}.hpp
//   include/boundary/reflecting_walls.hpp
//   include/backend/serial/reconstruction/{muscl,piecewise_const}.hpp
//   include/backend/serial/riemann/{HLL,rusanov,ROE}.hpp
//   include/backend/serial/flux_assembly/{flux_assembly,flux_assembly_bathy}.hpp
//   include/backend/serial/finite_volume/{finite_volume,finite_volume_bathy}.hpp
//   include/backend/serial/ssp_rk3/ssp_rk3.hpp
//   include/solver_assembly/time_utils.hpp
// What this file adds is only the glue that SerialSolver keeps in private members of a
// class that cannot be compiled without NetCDF (include/io/netCDF_writer.hpp:8): the
// compute_rhs / compute_rhs_bathy sequences (serial_solver.hpp:306-356), the step loop
// (:180-300) and enforce_positivity (:453-472), restated over raw arrays.  g is the
// reference's compile-time 9.80665.
#include "include/backend/serial/finite_volume/finite_volume.hpp"
#include "include/backend/serial/finite_volume/finite_volume_bathy.hpp"
#include "include/backend/serial/flux_assembly/flux_assembly.hpp"
#include "include/backend/serial/flux_assembly/flux_assembly_bathy.hpp"
#include "include/backend/serial/reconstruction/muscl.hpp"
#include "include/backend/serial/reconstruction/piecewise_const.hpp"
#include "include/backend/serial/riemann/HLL.hpp"
#include "include/backend/serial/riemann/ROE.hpp"
#include "include/backend/serial/riemann/rusanov.hpp"
#include "include/backend/serial/ssp_rk3/ssp_rk3.hpp"
#include "include/boundary/reflecting_walls.hpp"
#include "include/solver_assembly/time_utils.hpp"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>

namespace {

void load(State &U, const double *h, const double *hu, const double *hv) {
    const std::size_t n = U.h().size();
    std::memcpy(U.h().data(), h, n * sizeof(double));
    std::memcpy(U.hu().data(), hu, n * sizeof(double));
    std::memcpy(U.hv().data(), hv, n * sizeof(double));
}

void store(const State &U, double *h, double *hu, double *hv) {
    const std::size_t n = U.h().size();
    std::memcpy(h, U.h().data(), n * sizeof(double));
    std::memcpy(hu, U.hu().data(), n * sizeof(double));
    std::memcpy(hv, U.hv().data(), n * sizeof(double));
}

struct Setup {
    Grid grid;
    Array2D B;
    ReflectingWalls bc;
    Setup(int nx, int ny, double lx, double ly, const double *b)
        : grid(nx, ny, lx, ly, 2), B(grid.Nx_total(), grid.Ny_total()), bc(grid) {
        if (b) std::memcpy(B.data(), b, B.size() * sizeof(double));
    }
};

// serial_solver.hpp:306-356: ghost fill of U, fluxes, spatial operator
template <class Recon, class Riemann>
void rhs(Setup &S, bool bathy, State &U, State &L) {
    Recon recon;
    Riemann riemann;
    S.bc.apply_BC(U);
    if (bathy) {
        FluxAssemblyBathy fa;
        XFluxField xm(S.grid), xp(S.grid);
        YFluxField ym(S.grid), yp(S.grid);
        fa.compute_x_fluxes(U, S.B, recon, riemann, xm, xp, S.grid);
        fa.compute_y_fluxes(U, S.B, recon, riemann, ym, yp, S.grid);
        FiniteVolumeBathy(S.grid).apply_spatial_operator(L, xm, xp, ym, yp);
    } else {
        FluxAssembly fa;
        XFluxField x(S.grid);
        YFluxField y(S.grid);
        fa.compute_x_fluxes(U, recon, riemann, x, S.grid);
        fa.compute_y_fluxes(U, recon, riemann, y, S.grid);
        FiniteVolume(S.grid).apply_spatial_operator(L, x, y);
    }
}

// serial_solver.hpp:453-472
void positivity(const Grid &grid, State &U) {
    const int nG = grid.nG();
    for (int i = nG; i < nG + grid.Nx(); ++i)
        for (int j = nG; j < nG + grid.Ny(); ++j) {
            if (U.h()(i, j) < 1e-8) {
                U.h()(i, j) = 1e-8;
                U.hu()(i, j) = 0.0;
                U.hv()(i, j) = 0.0;
            }
            if (!std::isfinite(U.h()(i, j)) || !std::isfinite(U.hu()(i, j)) ||
                !std::isfinite(U.hv()(i, j)))
                throw std::runtime_error("non-finite state detected in positivity correction");
        }
}

// serial_solver.hpp:180-300
template <class Recon, class Riemann>
void run(Setup &S, bool bathy, State &U, int64_t nsteps, double dt) {
    State U1(S.grid), U2(S.grid), Un(S.grid), L(S.grid);
    SSPRK3 ti(dt);
    for (int64_t step = 0; step < nsteps; ++step) {
        rhs<Recon, Riemann>(S, bathy, U, L);
        ti.compute_U1(U1, U, L);
        if (bathy) positivity(S.grid, U1);
        rhs<Recon, Riemann>(S, bathy, U1, L);
        ti.compute_U2(U2, U1, U, L);
        if (bathy) positivity(S.grid, U2);
        rhs<Recon, Riemann>(S, bathy, U2, L);
        ti.compute_U_next(Un, U2, U, L);
        if (bathy) positivity(S.grid, Un);
        std::swap(U, Un);
        S.bc.apply_BC(U);
    }
}

template <class F>
int dispatch(int recon, int riemann, F &&f) {
    // recon: 0 MUSCL, 1 PiecewiseConst; riemann: 0 HLL, 1 Rusanov, 2 ROE
    if (recon == 0 && riemann == 0) return f(MUSCL{}, HLL{});
    if (recon == 0 && riemann == 1) return f(MUSCL{}, Rusanov{});
    if (recon == 0 && riemann == 2) return f(MUSCL{}, ROE{});
    if (recon == 1 && riemann == 0) return f(PiecewiseConst{}, HLL{});
    if (recon == 1 && riemann == 1) return f(PiecewiseConst{}, Rusanov{});
    if (recon == 1 && riemann == 2) return f(PiecewiseConst{}, ROE{});
    return -1;
}

}  // namespace

extern "C" {

// Returns 0, or 1 where the reference throws (HLL without dissipation speed, non-finite).
int ref_mod_rhs(int nx, int ny, double lx, double ly, int bathy, int recon, int riemann,
                double *h, double *hu, double *hv, const double *B, double *rh, double *rhu,
                double *rhv) {
    try {
        Setup S(nx, ny, lx, ly, B);
        State U(S.grid), L(S.grid);
        load(U, h, hu, hv);
        const int rc = dispatch(recon, riemann, [&](auto r, auto s) {
            rhs<decltype(r), decltype(s)>(S, bathy != 0, U, L);
            return 0;
        });
        store(U, h, hu, hv);
        store(L, rh, rhu, rhv);
        return rc;
    } catch (const std::exception &) {
        return 1;
    }
}

int ref_mod_run(int nx, int ny, double lx, double ly, int bathy, int recon, int riemann,
                double *h, double *hu, double *hv, const double *B, long nsteps, double dt) {
    try {
        Setup S(nx, ny, lx, ly, B);
        State U(S.grid);
        load(U, h, hu, hv);
        int rc = 0;
        try {
            rc = dispatch(recon, riemann, [&](auto r, auto s) {
                run<decltype(r), decltype(s)>(S, bathy != 0, U, nsteps, dt);
                return 0;
            });
        } catch (const std::exception &) {
            rc = 1;
        }
        store(U, h, hu, hv);
        return rc;
    } catch (const std::exception &) {
        return 1;
    }
}

// time_utils.hpp:13-36; -1 where it throws.
double ref_mod_stable_dt(int nx, int ny, double lx, double ly, const double *h, const double *hu,
                         const double *hv, double cfl) {
    try {
        Setup S(nx, ny, lx, ly, nullptr);
        State U(S.grid);
        load(U, h, hu, hv);
        return compute_stable_dt(U, S.grid, cfl);
    } catch (const std::exception &) {
        return -1.0;
    }
}

// One Riemann flux; returns 1 where HLL throws.
int ref_mod_riemann(int y, int kind, const double *L, const double *R, double *F) {
    try {
        const CellState l(L[0], L[1], L[2]), r(R[0], R[1], R[2]);
        CellState f;
        auto go = [&](const auto &s) { f = y ? s.y_flux(l, r) : s.x_flux(l, r); };
        if (kind == 0) go(HLL{});
        else if (kind == 1) go(Rusanov{});
        else go(ROE{});
        F[0] = f.h();
        F[1] = f.hu();
        F[2] = f.hv();
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

}  // extern "C"
