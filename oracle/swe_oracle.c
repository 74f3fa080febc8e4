/*
 * swe_oracle.c -- CPU restatement of the reference's optimized shallow-water step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP backend is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or as the timed CPU baseline -- never as a code
 * path of the product.  The product (libswe_hip.so) does not link it.
 *
 * What it restates (all paths relative to the reference tree):
 *   include/backend/openmp/omp_kernels.hpp  (whole file; HPC twin is
 *                                            include/backend/serial/HPC/HPC_kernels.hpp)
 *   include/backend/openmp/omp_solver.hpp:250-338  rk3_step
 *   include/backend/openmp/omp_solver.hpp:340-432  reflecting ghost fill
 *   include/constants.hpp:4-5                      g, eps
 *
 * Arithmetic order follows the reference expression by expression so results are
 * bitwise equal when both are built without FP contraction (-ffp-contract=off, no
 * -march flags that enable FMA).  g is a run-time parameter (the reference fixes it
 * at 9.80665 as a constexpr; folding 0.5*g at compile time or multiplying at run time
 * gives the same double).
 *
 * Extensions that the reference does not have, clearly separated:
 *   - bc_kind = SWE_BC_TRANSMISSIVE: zero-gradient ghost copy (no reference code).
 *   - adaptive dt: dt recomputed from compute_stable_dt before every step (the
 *     reference computes it once, at t = 0, omp_solver.hpp:151-162).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - bitwise against oracle/_ref (the reference's own omp_kernels.hpp compiled here)
 *     and against the golden fixtures in tests/golden/ made from it;
 *   - closed forms / invariants: lake at rest, zero wall mass flux, x<->y transpose
 *     symmetry, mass telescoping, HLL consistency, stable-dt closed form at rest.
 *   The transmissive ghost fill has no reference counterpart; it is pinned by its
 *   closed form (ghost == nearest interior cell) only.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define SWE_EPS 1.0e-8       /* include/constants.hpp:5 */
#define SWE_H_FLOOR 1.0e-8   /* omp_kernels.hpp:310 default argument */

enum { SWE_BC_REFLECTING = 0, SWE_BC_TRANSMISSIVE = 1 };

typedef struct {
    int32_t nx, ny, ng;   /* interior cells and ghost width */
    double dx, dy;        /* cell sizes (Lx/Nx, Ly/Ny as in include/core/grid.hpp:19-20) */
    double g;             /* gravity */
    int32_t bc_kind;      /* SWE_BC_* */
} oracle_grid;

/* ---- scalar helpers: omp_kernels.hpp:33-48 ---- */

static inline long idx_of(int i, int j, int stride) { return (long)i * stride + j; }

static inline double positive_height(double h) { return h > 0.0 ? h : 0.0; }

static inline double safe_velocity(double q, double h) { return (h > SWE_EPS) ? (q / h) : 0.0; }

/* std::min / std::max semantics: min(a,b) = (b < a) ? b : a ; max(a,b) = (a < b) ? b : a */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

double oracle_minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (a > 0.0 ? 1.0 : -1.0) * smin(fabs(a), fabs(b));
}

/* ---- HLL: omp_kernels.hpp:50-86 (x) and :88-124 (y) ---- */

void oracle_hll_x(double g, double hL, double huL, double hvL, double hR, double huR,
                  double hvR, double *out) {
    const double uL = safe_velocity(huL, hL);
    const double vL = safe_velocity(hvL, hL);
    const double uR = safe_velocity(huR, hR);
    const double vR = safe_velocity(hvR, hR);
    const double cL = sqrt(g * hL);
    const double cR = sqrt(g * hR);
    const double sL = smin(uL - cL, uR - cR);
    const double sR = smax(uL + cL, uR + cR);
    const double FL_h = huL;
    const double FL_hu = huL * uL + 0.5 * g * hL * hL;
    const double FL_hv = huL * vL;
    const double FR_h = huR;
    const double FR_hu = huR * uR + 0.5 * g * hR * hR;
    const double FR_hv = huR * vR;
    if (sL >= 0.0) {
        out[0] = FL_h; out[1] = FL_hu; out[2] = FL_hv;
    } else if (sR <= 0.0) {
        out[0] = FR_h; out[1] = FR_hu; out[2] = FR_hv;
    } else {
        const double inv = 1.0 / (sR - sL);
        out[0] = (sR * FL_h - sL * FR_h + sL * sR * (hR - hL)) * inv;
        out[1] = (sR * FL_hu - sL * FR_hu + sL * sR * (huR - huL)) * inv;
        out[2] = (sR * FL_hv - sL * FR_hv + sL * sR * (hvR - hvL)) * inv;
    }
}

void oracle_hll_y(double g, double hB, double huB, double hvB, double hT, double huT,
                  double hvT, double *out) {
    const double uB = safe_velocity(huB, hB);
    const double vB = safe_velocity(hvB, hB);
    const double uT = safe_velocity(huT, hT);
    const double vT = safe_velocity(hvT, hT);
    const double cB = sqrt(g * hB);
    const double cT = sqrt(g * hT);
    const double sB = smin(vB - cB, vT - cT);
    const double sT = smax(vB + cB, vT + cT);
    const double GB_h = hvB;
    const double GB_hu = hvB * uB;
    const double GB_hv = hvB * vB + 0.5 * g * hB * hB;
    const double GT_h = hvT;
    const double GT_hu = hvT * uT;
    const double GT_hv = hvT * vT + 0.5 * g * hT * hT;
    if (sB >= 0.0) {
        out[0] = GB_h; out[1] = GB_hu; out[2] = GB_hv;
    } else if (sT <= 0.0) {
        out[0] = GT_h; out[1] = GT_hu; out[2] = GT_hv;
    } else {
        const double inv = 1.0 / (sT - sB);
        out[0] = (sT * GB_h - sB * GT_h + sB * sT * (hT - hB)) * inv;
        out[1] = (sT * GB_hu - sB * GT_hu + sB * sT * (huT - huB)) * inv;
        out[2] = (sT * GB_hv - sB * GT_hv + sB * sT * (hvT - hvB)) * inv;
    }
}

/*
 * One face, reconstructed along a line (omp_kernels.hpp:141-189 for x, :209-257 for y).
 * The four cells are c0..c3 = (i-1, i, i+1, i+2) along the face normal; the face sits
 * between c1 and c2.  `normal_is_u` selects hll_x (x faces) or hll_y (y faces) and
 * which momentum gets the hydrostatic pressure correction.
 * out[0..2] = minus-side flux (h, hu, hv), out[3..5] = plus-side flux.
 */
static inline void face_flux(double g, int normal_is_u, const double *h, const double *hu,
                             const double *hv, const double *B, long k, long step,
                             double *out) {
    const long km = k - step, kr = k + step, krr = kr + step;
    const double eta_im1 = h[km] + B[km];
    const double eta_i = h[k] + B[k];
    const double eta_ip1 = h[kr] + B[kr];
    const double eta_ip2 = h[krr] + B[krr];

    const double eta_minus = eta_i + 0.5 * oracle_minmod(eta_i - eta_im1, eta_ip1 - eta_i);
    const double eta_plus = eta_ip1 - 0.5 * oracle_minmod(eta_ip1 - eta_i, eta_ip2 - eta_ip1);

    const double hu_minus = hu[k] + 0.5 * oracle_minmod(hu[k] - hu[km], hu[kr] - hu[k]);
    const double hv_minus = hv[k] + 0.5 * oracle_minmod(hv[k] - hv[km], hv[kr] - hv[k]);
    const double hu_plus = hu[kr] - 0.5 * oracle_minmod(hu[kr] - hu[k], hu[krr] - hu[kr]);
    const double hv_plus = hv[kr] - 0.5 * oracle_minmod(hv[kr] - hv[k], hv[krr] - hv[kr]);

    const double b_minus = B[k];
    const double b_plus = B[kr];
    const double b_star = smax(b_minus, b_plus);

    const double h_minus = positive_height(eta_minus - b_minus);
    const double h_plus = positive_height(eta_plus - b_plus);

    const double u_minus = safe_velocity(hu_minus, h_minus);
    const double v_minus = safe_velocity(hv_minus, h_minus);
    const double u_plus = safe_velocity(hu_plus, h_plus);
    const double v_plus = safe_velocity(hv_plus, h_plus);

    const double hs_minus = positive_height(eta_minus - b_star);
    const double hs_plus = positive_height(eta_plus - b_star);

    const double hL = hs_minus;
    const double huL = hs_minus * u_minus;
    const double hvL = hs_minus * v_minus;
    const double hR = hs_plus;
    const double huR = hs_plus * u_plus;
    const double hvR = hs_plus * v_plus;

    double f[3];
    if (normal_is_u)
        oracle_hll_x(g, hL, huL, hvL, hR, huR, hvR, f);
    else
        oracle_hll_y(g, hL, huL, hvL, hR, huR, hvR, f);

    const double corr_minus = 0.5 * g * (h_minus * h_minus - hs_minus * hs_minus);
    const double corr_plus = 0.5 * g * (h_plus * h_plus - hs_plus * hs_plus);

    out[0] = f[0];
    out[1] = normal_is_u ? f[1] + corr_minus : f[1];
    out[2] = normal_is_u ? f[2] : f[2] + corr_minus;
    out[3] = f[0];
    out[4] = normal_is_u ? f[1] + corr_plus : f[1];
    out[5] = normal_is_u ? f[2] : f[2] + corr_plus;
}

/* omp_kernels.hpp:126-192.  fm/fp are arrays of 3 pointers (h, hu, hv components). */
void oracle_x_fluxes(const oracle_grid *gr, const double *h, const double *hu, const double *hv,
                     const double *B, double *fm_h, double *fm_hu, double *fm_hv, double *fp_h,
                     double *fp_hu, double *fp_hv) {
    const int nG = gr->ng, s = gr->ny + 2 * gr->ng;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = nG - 1; i < nG + gr->nx; ++i) {
        for (int j = nG; j < nG + gr->ny; ++j) {
            const long k = idx_of(i, j, s);
            double o[6];
            face_flux(gr->g, 1, h, hu, hv, B, k, s, o);
            fm_h[k] = o[0]; fm_hu[k] = o[1]; fm_hv[k] = o[2];
            fp_h[k] = o[3]; fp_hu[k] = o[4]; fp_hv[k] = o[5];
        }
    }
}

/* omp_kernels.hpp:194-260 */
void oracle_y_fluxes(const oracle_grid *gr, const double *h, const double *hu, const double *hv,
                     const double *B, double *fm_h, double *fm_hu, double *fm_hv, double *fp_h,
                     double *fp_hu, double *fp_hv) {
    const int nG = gr->ng, s = gr->ny + 2 * gr->ng;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = nG; i < nG + gr->nx; ++i) {
        for (int j = nG - 1; j < nG + gr->ny; ++j) {
            const long k = idx_of(i, j, s);
            double o[6];
            face_flux(gr->g, 0, h, hu, hv, B, k, 1, o);
            fm_h[k] = o[0]; fm_hu[k] = o[1]; fm_hv[k] = o[2];
            fp_h[k] = o[3]; fp_hu[k] = o[4]; fp_hv[k] = o[5];
        }
    }
}

/* omp_kernels.hpp:262-286.  f points at 12 arrays: fxm(h,hu,hv) fxp(..) fym(..) fyp(..). */
void oracle_divergence(const oracle_grid *gr, double *const *f, double *rhs_h, double *rhs_hu,
                       double *rhs_hv) {
    const int nG = gr->ng, s = gr->ny + 2 * gr->ng;
    const double inv_dx = 1.0 / gr->dx;
    const double inv_dy = 1.0 / gr->dy;
    double *const *fxm = f, *const *fxp = f + 3, *const *fym = f + 6, *const *fyp = f + 9;
    double *rhs[3] = {rhs_h, rhs_hu, rhs_hv};
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = nG; i < nG + gr->nx; ++i) {
        for (int j = nG; j < nG + gr->ny; ++j) {
            const long k = idx_of(i, j, s);
            const long kmx = k - s, kmy = k - 1;
            for (int c = 0; c < 3; ++c)
                rhs[c][k] = -inv_dx * (fxm[c][k] - fxp[c][kmx]) - inv_dy * (fym[c][k] - fyp[c][kmy]);
        }
    }
}

/* omp_kernels.hpp:288-307.  Also returns the max wave speed through *max_speed_out. */
double oracle_stable_dt(const oracle_grid *gr, const double *h, const double *hu, const double *hv,
                        double cfl, double *max_speed_out) {
    const int nG = gr->ng, s = gr->ny + 2 * gr->ng;
    const double g = gr->g;
    double max_speed = 0.0;
#pragma omp parallel for collapse(2) schedule(static) reduction(max : max_speed)
    for (int i = nG; i < nG + gr->nx; ++i) {
        for (int j = nG; j < nG + gr->ny; ++j) {
            const long k = idx_of(i, j, s);
            const double hk = h[k];
            const double u = safe_velocity(hu[k], hk);
            const double v = safe_velocity(hv[k], hk);
            const double c = sqrt(g * smax(0.0, hk));
            max_speed = smax(max_speed, fabs(u) + c);
            max_speed = smax(max_speed, fabs(v) + c);
        }
    }
    if (max_speed_out) *max_speed_out = max_speed;
    if (max_speed <= SWE_EPS) return 1.0e100;
    return cfl * smin(gr->dx, gr->dy) / max_speed;
}

/* omp_kernels.hpp:309-331.  Returns 1 if any interior value is non-finite after the floor. */
int oracle_enforce_positivity(const oracle_grid *gr, double *h, double *hu, double *hv) {
    const int nG = gr->ng, s = gr->ny + 2 * gr->ng;
    int nonfinite = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(max : nonfinite)
    for (int i = nG; i < nG + gr->nx; ++i) {
        for (int j = nG; j < nG + gr->ny; ++j) {
            const long k = idx_of(i, j, s);
            if (h[k] < SWE_H_FLOOR) {
                h[k] = SWE_H_FLOOR;
                hu[k] = 0.0;
                hv[k] = 0.0;
            }
            if (!isfinite(h[k]) || !isfinite(hu[k]) || !isfinite(hv[k])) nonfinite = 1;
        }
    }
    return nonfinite;
}

/*
 * Ghost fill.  Reflecting: omp_solver.hpp:340-393 -- x walls over interior j with hu
 * negated, then y walls over every i (corners included) with hv negated.
 * Transmissive (extension): the same two passes, copying the nearest interior cell.
 * A NULL hu/hv makes it the scalar fill of omp_solver.hpp:395-432 (no sign flips).
 */
void oracle_apply_bc(const oracle_grid *gr, double *h, double *hu, double *hv) {
    const int nG = gr->ng, Nx = gr->nx, Ny = gr->ny, s = gr->ny + 2 * gr->ng;
    const int nxt = gr->nx + 2 * gr->ng;
    const int refl = gr->bc_kind == SWE_BC_REFLECTING;
    for (int gg = 0; gg < nG; ++gg) {
        const int iL = nG - 1 - gg, iR = nG + Nx + gg;
        const int iSrcL = refl ? nG + gg : nG, iSrcR = refl ? nG + Nx - 1 - gg : nG + Nx - 1;
        for (int j = nG; j < nG + Ny; ++j) {
            const long kL = idx_of(iL, j, s), kSL = idx_of(iSrcL, j, s);
            const long kR = idx_of(iR, j, s), kSR = idx_of(iSrcR, j, s);
            h[kL] = h[kSL];
            h[kR] = h[kSR];
            if (hu) {
                hu[kL] = refl ? -hu[kSL] : hu[kSL];
                hu[kR] = refl ? -hu[kSR] : hu[kSR];
                hv[kL] = hv[kSL];
                hv[kR] = hv[kSR];
            }
        }
    }
    for (int gg = 0; gg < nG; ++gg) {
        const int jB = nG - 1 - gg, jT = nG + Ny + gg;
        const int jSrcB = refl ? nG + gg : nG, jSrcT = refl ? nG + Ny - 1 - gg : nG + Ny - 1;
        for (int i = 0; i < nxt; ++i) {
            const long kB = idx_of(i, jB, s), kSB = idx_of(i, jSrcB, s);
            const long kT = idx_of(i, jT, s), kST = idx_of(i, jSrcT, s);
            h[kB] = h[kSB];
            h[kT] = h[kST];
            if (hu) {
                hu[kB] = hu[kSB];
                hu[kT] = hu[kST];
                hv[kB] = refl ? -hv[kSB] : hv[kSB];
                hv[kT] = refl ? -hv[kST] : hv[kST];
            }
        }
    }
}

/* ---- driver: omp_solver.hpp:220-248 (compute_rhs) and :250-338 (rk3_step) ---- */

typedef struct {
    oracle_grid gr;
    long n;
    double *B;
    double *h1, *hu1, *hv1, *h2, *hu2, *hv2;
    double *rhs[3];
    double *f[12];
} oracle_ws;

static void ws_free(oracle_ws *w) {
    free(w->h1); free(w->hu1); free(w->hv1);
    free(w->h2); free(w->hu2); free(w->hv2);
    for (int c = 0; c < 3; ++c) free(w->rhs[c]);
    for (int c = 0; c < 12; ++c) free(w->f[c]);
}

static int ws_init(oracle_ws *w, const oracle_grid *gr, const double *B) {
    memset(w, 0, sizeof(*w));
    w->gr = *gr;
    w->n = (long)(gr->nx + 2 * gr->ng) * (gr->ny + 2 * gr->ng);
    const size_t bytes = (size_t)w->n * sizeof(double);
    double **all[] = {&w->h1, &w->hu1, &w->hv1, &w->h2, &w->hu2, &w->hv2,
                      &w->rhs[0], &w->rhs[1], &w->rhs[2]};
    for (size_t a = 0; a < sizeof(all) / sizeof(all[0]); ++a)
        if (!(*all[a] = (double *)calloc(1, bytes))) { ws_free(w); return -1; }
    for (int c = 0; c < 12; ++c)
        if (!(w->f[c] = (double *)calloc(1, bytes))) { ws_free(w); return -1; }
    w->B = (double *)B;
    return 0;
}

static void compute_rhs(oracle_ws *w, const double *h, const double *hu, const double *hv) {
    oracle_x_fluxes(&w->gr, h, hu, hv, w->B, w->f[0], w->f[1], w->f[2], w->f[3], w->f[4], w->f[5]);
    oracle_y_fluxes(&w->gr, h, hu, hv, w->B, w->f[6], w->f[7], w->f[8], w->f[9], w->f[10], w->f[11]);
    oracle_divergence(&w->gr, w->f, w->rhs[0], w->rhs[1], w->rhs[2]);
}

/* Returns 0, or the stage (1..3) whose output was non-finite: the reference throws there
 * (omp_kernels.hpp:322-330), leaving U^n untouched after a stage-1 or stage-2 failure and the
 * floored stage-3 output in U after a stage-3 failure (omp_solver.hpp:250-338). */
static int rk3_step(oracle_ws *w, double *h, double *hu, double *hv, double dt) {
    const long n = w->n;
    double *r0 = w->rhs[0], *r1 = w->rhs[1], *r2 = w->rhs[2];
    int bad = 0;
    oracle_apply_bc(&w->gr, h, hu, hv);
    compute_rhs(w, h, hu, hv);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {
        w->h1[k] = h[k] + dt * r0[k];
        w->hu1[k] = hu[k] + dt * r1[k];
        w->hv1[k] = hv[k] + dt * r2[k];
    }
    bad |= oracle_enforce_positivity(&w->gr, w->h1, w->hu1, w->hv1);
    if (bad) return 1;  /* stage 1 */
    oracle_apply_bc(&w->gr, w->h1, w->hu1, w->hv1);
    compute_rhs(w, w->h1, w->hu1, w->hv1);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {
        w->h2[k] = 0.75 * h[k] + 0.25 * (w->h1[k] + dt * r0[k]);
        w->hu2[k] = 0.75 * hu[k] + 0.25 * (w->hu1[k] + dt * r1[k]);
        w->hv2[k] = 0.75 * hv[k] + 0.25 * (w->hv1[k] + dt * r2[k]);
    }
    bad |= oracle_enforce_positivity(&w->gr, w->h2, w->hu2, w->hv2);
    if (bad) return 2;
    oracle_apply_bc(&w->gr, w->h2, w->hu2, w->hv2);
    compute_rhs(w, w->h2, w->hu2, w->hv2);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {
        h[k] = (1.0 / 3.0) * h[k] + (2.0 / 3.0) * (w->h2[k] + dt * r0[k]);
        hu[k] = (1.0 / 3.0) * hu[k] + (2.0 / 3.0) * (w->hu2[k] + dt * r1[k]);
        hv[k] = (1.0 / 3.0) * hv[k] + (2.0 / 3.0) * (w->hv2[k] + dt * r2[k]);
    }
    bad |= oracle_enforce_positivity(&w->gr, h, hu, hv);
    if (bad) return 3;
    oracle_apply_bc(&w->gr, h, hu, hv);
    return 0;
}

/* One compute_rhs (after a ghost fill of h/hu/hv in place); rhs written at interior cells. */
int oracle_rhs(const oracle_grid *gr, double *h, double *hu, double *hv, const double *B,
               double *rhs_h, double *rhs_hu, double *rhs_hv) {
    oracle_ws w;
    if (ws_init(&w, gr, B)) return -1;
    oracle_apply_bc(gr, h, hu, hv);
    compute_rhs(&w, h, hu, hv);
    const size_t bytes = (size_t)w.n * sizeof(double);
    memcpy(rhs_h, w.rhs[0], bytes);
    memcpy(rhs_hu, w.rhs[1], bytes);
    memcpy(rhs_hv, w.rhs[2], bytes);
    ws_free(&w);
    return 0;
}

/*
 * Advance nsteps SSP-RK3 steps in place.  B must already hold its ghosts (the caller
 * applies oracle_apply_bc(gr, B, NULL, NULL) as the reference does at setup).
 *   dt_mode 0: fixed dt = dt_or_cfl   (reference behaviour, omp_solver.hpp:99)
 *   dt_mode 1: adaptive, dt = compute_stable_dt(U^n, cfl = dt_or_cfl) before each step
 * dt_hist (optional, nsteps entries) receives the dt used by each step.
 * Returns 0, 1 on non-finite state (steps_done reports progress: the steps completed before
 * the failing one; failed_stage, optional, the stage 1..3 that failed, else 0), -1 on
 * allocation failure.
 */
int oracle_run(const oracle_grid *gr, double *h, double *hu, double *hv, const double *B,
               int dt_mode, double dt_or_cfl, long nsteps, double *dt_hist, long *steps_done,
               int *failed_stage) {
    oracle_ws w;
    if (ws_init(&w, gr, B)) return -1;
    int rc = 0;
    long s = 0;
    for (; s < nsteps; ++s) {
        double dt = dt_or_cfl;
        if (dt_mode == 1) dt = oracle_stable_dt(gr, h, hu, hv, dt_or_cfl, NULL);
        if (dt_hist) dt_hist[s] = dt;
        const int st = rk3_step(&w, h, hu, hv, dt);
        if (st) {
            if (failed_stage) *failed_stage = st;
            rc = 1;
            break;
        }
    }
    if (steps_done) *steps_done = s;
    if (failed_stage && rc == 0) *failed_stage = 0;
    ws_free(&w);
    return rc;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
