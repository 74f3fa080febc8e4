/*
 * swe_modular.c -- CPU restatement of the reference's modular serial solver
 * (SerialSolver<ReflectingWalls, Recon, Riemann, SSPRK3, Bathymetry>).
 *
 * TEST INFRASTRUCTURE ONLY, like swe_oracle.c: the checker of the HIP backend's
 * modular scheme variants.  Only tests/ may load it.  The product does not link it.
 *
 * What it restates (paths relative to the reference tree):
 *   include/core/cell_state.hpp:19-26,31-49         CellState::u/v and arithmetic
 *   include/backend/serial/riemann/riemann.hpp:12-19  physical fluxes F and G
 *   include/backend/serial/riemann/HLL.hpp:9-45       HLL
 *   include/backend/serial/riemann/rusanov.hpp:11-30  Rusanov
 *   include/backend/serial/riemann/ROE.hpp:9-70       ROE
 *   include/backend/serial/reconstruction/{piecewise_const,muscl}.hpp
 *   include/backend/serial/flux_assembly/flux_assembly.hpp:14-45       (no bathymetry)
 *   include/backend/serial/flux_assembly/flux_assembly_bathy.hpp:17-170 (hydrostatic)
 *   include/backend/serial/finite_volume/{finite_volume,finite_volume_bathy}.hpp
 *   include/backend/serial/ssp_rk3/ssp_rk3.hpp:10-40  stage combinations
 *   include/backend/serial/serial_solver.hpp:180-300,453-478  step loop, positivity
 *   include/boundary/reflecting_walls.hpp:9-46        ghost fill (y walls first)
 *   include/solver_assembly/time_utils.hpp:13-36      CFL check formula
 *
 * Every expression keeps the reference's operand order (C evaluates left to right like
 * the reference's C++ operators), so results are bitwise equal when both sides are built
 * with -ffp-contract=off and no FMA-enabling -march.  g is a run-time parameter.
 *
 * Pins (tests/test_oracle_modular.py): bitwise against oracle/_ref's build of the
 * reference's own modular headers; and closed forms: lake at rest (flat, no bathymetry:
 * rhs == 0), consistency F(U, U) = physical flux for every Riemann solver, x<->y
 * transpose symmetry of the whole step, mass telescoping.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MOD_EPS 1.0e-8     /* include/constants.hpp:5 */
#define MOD_FLOOR 1.0e-8   /* serial_solver.hpp:454 */

enum { MOD_RECON_MUSCL = 0, MOD_RECON_PC = 1 };
enum { MOD_RIEMANN_HLL = 0, MOD_RIEMANN_RUSANOV = 1, MOD_RIEMANN_ROE = 2 };
enum { MOD_OK = 0, MOD_NONFINITE = 1, MOD_HLL_FAIL = 2, MOD_BAD = 3 };

typedef struct {
    int32_t nx, ny, ng;
    double dx, dy;
    double g;
    int32_t bc_kind; /* reflecting only (ignored) */
} mod_grid;

typedef struct {
    double h, hu, hv;
} cs;

static inline double smin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static inline double smax(double a, double b) { return (a < b) ? b : a; } /* std::max */

/* cell_state.hpp:19-26 */
static inline double cs_u(cs U) { return (U.h > MOD_EPS) ? (U.hu / U.h) : 0.0; }
static inline double cs_v(cs U) { return (U.h > MOD_EPS) ? (U.hv / U.h) : 0.0; }

static inline cs cs_add(cs a, cs b) { return (cs){a.h + b.h, a.hu + b.hu, a.hv + b.hv}; }
static inline cs cs_sub(cs a, cs b) { return (cs){a.h - b.h, a.hu - b.hu, a.hv - b.hv}; }
static inline cs cs_scale(double s, cs a) { return (cs){s * a.h, s * a.hu, s * a.hv}; }
static inline cs cs_div(cs a, double s) { return (cs){a.h / s, a.hu / s, a.hv / s}; }

/* riemann.hpp:12-19 */
static inline cs phys_F(double g, cs U) {
    return (cs){U.hu, cs_u(U) * cs_u(U) * U.h + 0.5 * g * U.h * U.h, U.hu * cs_v(U)};
}
static inline cs phys_G(double g, cs U) {
    return (cs){U.hv, U.hu * cs_v(U), U.h * cs_v(U) * cs_v(U) + 0.5 * g * U.h * U.h};
}

/* HLL.hpp:9-45; y: the same with v and G.  *fail on the reference's throw. */
static cs hll(double g, int y, cs L, cs R, int *fail) {
    const double uL = y ? cs_v(L) : cs_u(L), uR = y ? cs_v(R) : cs_u(R);
    const double sR = smax(uL + sqrt(g * L.h), uR + sqrt(g * R.h));
    const double sL = smin(uL - sqrt(g * L.h), uR - sqrt(g * R.h));
    const cs FL = y ? phys_G(g, L) : phys_F(g, L);
    const cs FR = y ? phys_G(g, R) : phys_F(g, R);
    if (sL >= 0.0) return FL;
    if (sL < 0.0 && sR > 0.0) {
        const cs num = cs_add(cs_sub(cs_scale(sR, FL), cs_scale(sL, FR)),
                              cs_scale(sL * sR, cs_sub(R, L)));
        return cs_div(num, sR - sL);
    }
    if (sR <= 0.0) return FR;
    *fail = 1;
    return (cs){NAN, NAN, NAN};
}

/* rusanov.hpp:11-30 */
static cs rusanov(double g, int y, cs L, cs R) {
    const double uL = y ? cs_v(L) : cs_u(L), uR = y ? cs_v(R) : cs_u(R);
    const double hL = smax(L.h, 0.0), hR = smax(R.h, 0.0);
    const double a = smax(fabs(uL) + sqrt(g * hL), fabs(uR) + sqrt(g * hR));
    const cs FL = y ? phys_G(g, L) : phys_F(g, L);
    const cs FR = y ? phys_G(g, R) : phys_F(g, R);
    return cs_sub(cs_scale(0.5, cs_add(FL, FR)), cs_scale(0.5 * a, cs_sub(R, L)));
}

/* ROE.hpp:66-69 */
static inline double roe_avg(double u1, double h1, double u2, double h2) {
    return ((sqrt(h1) * u1) + (sqrt(h2) * u2)) / (sqrt(h1) + sqrt(h2));
}

/* ROE.hpp:9-61.  In (h, normal, tangential) terms both directions read the same; the
 * arrays r1, r2, r3 multiply as array[k] * scalar and add commutatively, as the
 * reference's array operators (cell_state.hpp:52-64) do. */
static cs roe(double g, int y, cs L, cs R) {
    const double h = (L.h + R.h) / 2.0;
    const double c = sqrt(g * h);
    const double u_roe = roe_avg(cs_u(L), L.h, cs_u(R), R.h);
    const double v_roe = roe_avg(cs_v(L), L.h, cs_v(R), R.h);
    const double un = y ? v_roe : u_roe, ut = y ? u_roe : v_roe;
    const double l1 = un - c, l2 = un, l3 = un + c;
    const cs dU = cs_sub(R, L);
    const double dn = y ? dU.hv : dU.hu, dt = y ? dU.hu : dU.hv;
    const double a1 = 0.5 * (dU.h - ((dn - un * dU.h) / c));
    const double a3 = 0.5 * (dU.h + ((dn - un * dU.h) / c));
    const double a2 = dt - ut * dU.h;
    const double s1 = fabs(l1) * a1, s2 = fabs(l2) * a2, s3 = fabs(l3) * a3;
    /* r1 = (1, l1, ut), r2 = (0, 0, 1), r3 = (1, l3, ut) in (h, normal, tangential) */
    const double dh = ((1.0 * s1) + (0.0 * s2)) + (1.0 * s3);
    const double dnn = ((l1 * s1) + (0.0 * s2)) + (l3 * s3);
    const double dtt = ((ut * s1) + (1.0 * s2)) + (ut * s3);
    const cs FL = y ? phys_G(g, L) : phys_F(g, L);
    const cs FR = y ? phys_G(g, R) : phys_F(g, R);
    const cs avg = cs_scale(0.5, cs_add(FL, FR));
    const double hh = dh * 0.5, hn = dnn * 0.5, ht = dtt * 0.5;
    return y ? (cs){avg.h - hh, avg.hu - ht, avg.hv - hn} : (cs){avg.h - hh, avg.hu - hn, avg.hv - ht};
}

static cs riemann(double g, int y, int kind, cs L, cs R, int *fail) {
    if (kind == MOD_RIEMANN_RUSANOV) return rusanov(g, y, L, R);
    if (kind == MOD_RIEMANN_ROE) return roe(g, y, L, R);
    return hll(g, y, L, R, fail);
}

/* muscl.hpp:7-10 */
static inline double minmod_scalar(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (a > 0.0 ? 1.0 : -1.0) * smin(fabs(a), fabs(b));
}

static inline long kk(int i, int j, int s) { return (long)i * s + j; }

/* Cell (i, j) of a state W = (w, hu, hv) (w = h, or eta with bathymetry). */
static inline cs cell(const double *w, const double *hu, const double *hv, int i, int j, int s) {
    const long k = kk(i, j, s);
    return (cs){w[k], hu[k], hv[k]};
}

/* muscl.hpp:48-66: the limited slope of cell (i, j) along x (y = 0) or y (y = 1). */
static cs slope(const double *w, const double *hu, const double *hv, int i, int j, int s, int y) {
    const int di = y ? 0 : 1, dj = y ? 1 : 0;
    const cs c = cell(w, hu, hv, i, j, s);
    const cs m = cell(w, hu, hv, i - di, j - dj, s);
    const cs p = cell(w, hu, hv, i + di, j + dj, s);
    return (cs){minmod_scalar(c.h - m.h, p.h - c.h), minmod_scalar(c.hu - m.hu, p.hu - c.hu),
                minmod_scalar(c.hv - m.hv, p.hv - c.hv)};
}

/* The two reconstructed states of the face between cell (i, j) and its +x (y = 0) or +y
 * (y = 1) neighbour: piecewise_const.hpp:13-29 / muscl.hpp:14-41. */
static void recon(int kind, const double *w, const double *hu, const double *hv, int i, int j,
                  int s, int y, cs *m, cs *p) {
    const int di = y ? 0 : 1, dj = y ? 1 : 0;
    *m = cell(w, hu, hv, i, j, s);
    *p = cell(w, hu, hv, i + di, j + dj, s);
    if (kind == MOD_RECON_MUSCL) {
        *m = cs_add(*m, cs_scale(0.5, slope(w, hu, hv, i, j, s, y)));
        *p = cs_sub(*p, cs_scale(0.5, slope(w, hu, hv, i + di, j + dj, s, y)));
    }
}

/* flux_assembly_bathy.hpp:233-240 (note the test: h < eps gives 0, so h == eps divides). */
static inline double fa_velocity(double q, double h) { return (h < MOD_EPS) ? 0.0 : q / h; }
static inline double fa_positive(double h) { return smax(0.0, h); }

/* ---- reflecting_walls.hpp:9-46: y walls over every i, then x walls over every j ---- */
void oracle_mod_apply_bc(const mod_grid *gr, double *h, double *hu, double *hv) {
    const int nG = gr->ng, s = gr->ny + 2 * gr->ng;
    const int nxt = gr->nx + 2 * nG, nyt = s;
    for (int g = 0; g < nG; ++g) {
        const int jgb = nG - 1 - g, jib = nG + g, jgt = nyt - nG + g, jit = nyt - nG - 1 - g;
        for (int i = 0; i < nxt; ++i) {
            h[kk(i, jgb, s)] = h[kk(i, jib, s)];
            h[kk(i, jgt, s)] = h[kk(i, jit, s)];
            if (hu) {
                hu[kk(i, jgb, s)] = hu[kk(i, jib, s)];
                hv[kk(i, jgb, s)] = -hv[kk(i, jib, s)];
                hu[kk(i, jgt, s)] = hu[kk(i, jit, s)];
                hv[kk(i, jgt, s)] = -hv[kk(i, jit, s)];
            }
        }
    }
    for (int g = 0; g < nG; ++g) {
        const int igl = nG - 1 - g, iil = nG + g, igr = nxt - nG + g, iir = nxt - nG - 1 - g;
        for (int j = 0; j < nyt; ++j) {
            h[kk(igl, j, s)] = h[kk(iil, j, s)];
            h[kk(igr, j, s)] = h[kk(iir, j, s)];
            if (hu) {
                hu[kk(igl, j, s)] = -hu[kk(iil, j, s)];
                hv[kk(igl, j, s)] = hv[kk(iil, j, s)];
                hu[kk(igr, j, s)] = -hu[kk(iir, j, s)];
                hv[kk(igr, j, s)] = hv[kk(iir, j, s)];
            }
        }
    }
}

/*
 * rhs of a state (its ghosts are filled here first, as compute_rhs does:
 * serial_solver.hpp:306-328 / :330-356).  bathy = 0: flux_assembly.hpp on U;
 * bathy = 1: flux_assembly_bathy.hpp on W = (h + B, hu, hv) with the hydrostatic states.
 * Returns MOD_HLL_FAIL where the reference's HLL would throw.
 */
int oracle_mod_rhs(const mod_grid *gr, int bathy, int recon_kind, int riemann_kind, double *h,
                   double *hu, double *hv, const double *B, double *rh, double *rhu, double *rhv) {
    const int nG = gr->ng, Nx = gr->nx, Ny = gr->ny, s = Ny + 2 * nG;
    const long n = (long)(Nx + 2 * nG) * s;
    const double g = gr->g;
    oracle_mod_apply_bc(gr, h, hu, hv);
    double *w = h;
    double *wbuf = NULL;
    if (bathy) {
        wbuf = (double *)malloc(sizeof(double) * n);
        for (long k = 0; k < n; ++k) wbuf[k] = h[k] + B[k];
        w = wbuf;
    }
    /* face fluxes: minus/plus per face (equal without bathymetry) */
    double *fx[6], *fy[6];
    for (int c = 0; c < 6; ++c) {
        fx[c] = (double *)calloc(n + s, sizeof(double));
        fy[c] = (double *)calloc(n + s, sizeof(double));
    }
    int fail = 0;
    for (int y = 0; y <= 1; ++y) {
        const int i0 = y ? nG : nG - 1, i1 = nG + Nx;
        const int j0 = y ? nG - 1 : nG, j1 = nG + Ny;
        double **f = y ? fy : fx;
#pragma omp parallel for reduction(| : fail) schedule(static)
        for (int i = i0; i < i1; ++i) {
            for (int j = j0; j < j1; ++j) {
                cs m, p;
                recon(recon_kind, w, hu, hv, i, j, s, y, &m, &p);
                const long k = kk(i, j, s);
                int fl = 0;
                if (!bathy) {
                    const cs F = riemann(g, y, riemann_kind, m, p, &fl);
                    f[0][k] = f[3][k] = F.h;
                    f[1][k] = f[4][k] = F.hu;
                    f[2][k] = f[5][k] = F.hv;
                } else {
                    const double b_m = B[k], b_p = y ? B[kk(i, j + 1, s)] : B[kk(i + 1, j, s)];
                    const double b_star = smax(b_m, b_p);
                    const double h_m = fa_positive(m.h - b_m), h_p = fa_positive(p.h - b_p);
                    const double u_m = fa_velocity(m.hu, h_m), v_m = fa_velocity(m.hv, h_m);
                    const double u_p = fa_velocity(p.hu, h_p), v_p = fa_velocity(p.hv, h_p);
                    const double hs_m = fa_positive(m.h - b_star), hs_p = fa_positive(p.h - b_star);
                    const cs Um = {hs_m, hs_m * u_m, hs_m * v_m};
                    const cs Up = {hs_p, hs_p * u_p, hs_p * v_p};
                    const cs F = riemann(g, y, riemann_kind, Um, Up, &fl);
                    const double corr_m = 0.5 * g * (h_m * h_m - hs_m * hs_m);
                    const double corr_p = 0.5 * g * (h_p * h_p - hs_p * hs_p);
                    f[0][k] = F.h;
                    f[3][k] = F.h;
                    if (!y) {
                        f[1][k] = F.hu + corr_m;
                        f[4][k] = F.hu + corr_p;
                        f[2][k] = f[5][k] = F.hv;
                    } else {
                        f[1][k] = f[4][k] = F.hu;
                        f[2][k] = F.hv + corr_m;
                        f[5][k] = F.hv + corr_p;
                    }
                }
                fail |= fl;
            }
        }
    }
    /* finite_volume(_bathy).hpp: -1/dx * (F-[i] - F+[i-1]) - 1/dy * (G-[j] - G+[j-1]) */
    double *out[3] = {rh, rhu, rhv};
#pragma omp parallel for schedule(static)
    for (int i = nG; i < nG + Nx; ++i)
        for (int j = nG; j < nG + Ny; ++j) {
            const long k = kk(i, j, s), kx = kk(i - 1, j, s), ky = kk(i, j - 1, s);
            for (int c = 0; c < 3; ++c)
                out[c][k] = (-1.0 / gr->dx) * (fx[c][k] - fx[3 + c][kx]) -
                            (1.0 / gr->dy) * (fy[c][k] - fy[3 + c][ky]);
        }
    for (int c = 0; c < 6; ++c) {
        free(fx[c]);
        free(fy[c]);
    }
    free(wbuf);
    return fail ? MOD_HLL_FAIL : MOD_OK;
}

/* serial_solver.hpp:453-472 */
static int positivity(const mod_grid *gr, double *h, double *hu, double *hv) {
    const int nG = gr->ng, s = gr->ny + 2 * nG;
    int bad = 0;
    for (int i = nG; i < nG + gr->nx; ++i)
        for (int j = nG; j < nG + gr->ny; ++j) {
            const long k = kk(i, j, s);
            if (h[k] < MOD_FLOOR) {
                h[k] = MOD_FLOOR;
                hu[k] = 0.0;
                hv[k] = 0.0;
            }
            if (!isfinite(h[k]) || !isfinite(hu[k]) || !isfinite(hv[k])) bad = 1;
        }
    return bad;
}

/*
 * nsteps steps of the serial solver's loop (serial_solver.hpp:180-216 with bathymetry,
 * :255-300 without): per stage the rhs, ssp_rk3.hpp's combination over every array
 * element, and with bathymetry the positivity correction; after each step the ghost
 * fill.  Returns MOD_NONFINITE / MOD_HLL_FAIL at the first failing stage.
 */
int oracle_mod_run(const mod_grid *gr, int bathy, int recon_kind, int riemann_kind, double *h,
                   double *hu, double *hv, const double *B, int64_t nsteps, double dt) {
    const int s = gr->ny + 2 * gr->ng;
    const long n = (long)(gr->nx + 2 * gr->ng) * s;
    double *buf = (double *)calloc(9 * n, sizeof(double));
    double *U1[3] = {buf, buf + n, buf + 2 * n}, *U2[3] = {buf + 3 * n, buf + 4 * n, buf + 5 * n};
    double *L[3] = {buf + 6 * n, buf + 7 * n, buf + 8 * n};
    double *U[3] = {h, hu, hv};
    int rc = MOD_OK;
    const double c2 = dt / 4.0, c3 = 2.0 * dt / 3.0;
    for (int64_t step = 0; step < nsteps && rc == MOD_OK; ++step) {
        /* stage 1: U1 = 1.0*U + dt*L(U) */
        rc = oracle_mod_rhs(gr, bathy, recon_kind, riemann_kind, U[0], U[1], U[2], B, L[0], L[1], L[2]);
        if (rc) break;
        for (int c = 0; c < 3; ++c)
            for (long k = 0; k < n; ++k) U1[c][k] = 1.0 * U[c][k] + dt * L[c][k];
        if (bathy && positivity(gr, U1[0], U1[1], U1[2])) { rc = MOD_NONFINITE; break; }
        /* stage 2: U2 = 3/4*U + 1/4*U1 + dt/4*L(U1) */
        rc = oracle_mod_rhs(gr, bathy, recon_kind, riemann_kind, U1[0], U1[1], U1[2], B, L[0], L[1], L[2]);
        if (rc) break;
        for (int c = 0; c < 3; ++c)
            for (long k = 0; k < n; ++k)
                U2[c][k] = 3.0 / 4.0 * U[c][k] + 1.0 / 4.0 * U1[c][k] + c2 * L[c][k];
        if (bathy && positivity(gr, U2[0], U2[1], U2[2])) { rc = MOD_NONFINITE; break; }
        /* stage 3: U = 1/3*U + 2/3*U2 + 2dt/3*L(U2) */
        rc = oracle_mod_rhs(gr, bathy, recon_kind, riemann_kind, U2[0], U2[1], U2[2], B, L[0], L[1], L[2]);
        if (rc) break;
        for (int c = 0; c < 3; ++c)
            for (long k = 0; k < n; ++k)
                U[c][k] = 1.0 / 3.0 * U[c][k] + 2.0 / 3.0 * U2[c][k] + c3 * L[c][k];
        if (bathy && positivity(gr, U[0], U[1], U[2])) { rc = MOD_NONFINITE; break; }
        oracle_mod_apply_bc(gr, U[0], U[1], U[2]);
    }
    free(buf);
    return rc;
}

/* time_utils.hpp:13-36: min over interior cells of cfl / ((|u| + c)/dx + (|v| + c)/dy)
 * with u = hu/h (no eps test).  Returns -1 where the reference throws (h < 0). */
double oracle_mod_stable_dt(const mod_grid *gr, const double *h, const double *hu,
                            const double *hv, double cfl) {
    const int nG = gr->ng, s = gr->ny + 2 * nG;
    double lim = 1.7976931348623157e308;
    for (int i = nG; i < nG + gr->nx; ++i)
        for (int j = nG; j < nG + gr->ny; ++j) {
            const long k = kk(i, j, s);
            if (h[k] < 0.0) return -1.0;
            const double u = hu[k] / h[k], v = hv[k] / h[k];
            const double c = cfl / ((fabs(u) + sqrt(gr->g * h[k])) / gr->dx +
                                    (fabs(v) + sqrt(gr->g * h[k])) / gr->dy);
            lim = smin(lim, c);
        }
    return lim;
}

/* One face of the modular Riemann solvers, for scalar known-answer tests. */
void oracle_mod_riemann(double g, int y, int kind, const double *L, const double *R, double *F,
                        int *fail) {
    const cs l = {L[0], L[1], L[2]}, r = {R[0], R[1], R[2]};
    int fl = 0;
    const cs f = riemann(g, y, kind, l, r, &fl);
    F[0] = f.h;
    F[1] = f.hu;
    F[2] = f.hv;
    *fail = fl;
}
