// swe_tangent.cpp -- forward-mode (dual-number) tangent of the serial optimized solver.
//
// TEST INFRASTRUCTURE ONLY: the checker of the GPU discrete adjoint.  Never linked into
// the product.
//
// The scalar type T is double (primal only) or Dual (value + one directional derivative).
// The arithmetic restates the optimized serial solver, include/backend/serial/HPC/
// HPC_kernels.hpp (face fluxes :122-254, divergence :256-279, positivity :300-315) with the
// rk3_step sequence of include/backend/serial/HPC/HPC_solver.hpp:233-318 and the ghost fill
// of :320-404, in the same operation order as oracle/swe_oracle.c (so T = double is
// bitwise the oracle; tests check that).
//
// Derivative conventions at non-smooth points (the ones the GPU adjoint reverses):
//   minmod(a, b): 0 if a*b <= 0; otherwise the derivative of the argument std::min(|a|,|b|)
//                 selected (|a| on a tie);
//   positive_height(x) = x > 0 ? x : 0: derivative 1 iff x > 0;
//   std::min / std::max: derivative of the argument they return (first on a tie);
//   safe_velocity(q, h): 0 unless h > eps;
//   sqrt(x): derivative 0 at x = 0;
//   HLL: the branch taken by the primal values;
//   positivity floor: derivative 0 where the stage output is the floor state (1e-8, 0, 0);
//   dt: frozen (a given sequence, no derivative).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Dual {
    double v, d;
    Dual(double v_ = 0.0, double d_ = 0.0) : v(v_), d(d_) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
inline Dual operator*(double s, Dual a) { return Dual(s * a.v, s * a.d); }
inline double val(double x) { return x; }
inline double val(const Dual &x) { return x.v; }
inline double dsqrt(double x) { return std::sqrt(x); }
inline Dual dsqrt(Dual x) {
    const double r = std::sqrt(x.v);
    return Dual(r, r > 0.0 ? x.d / (2.0 * r) : 0.0);
}
inline double dabs(double x) { return std::fabs(x); }
inline Dual dabs(Dual x) { return x.v < 0.0 ? -x : (x.v > 0.0 ? x : Dual(std::fabs(x.v), 0.0)); }

constexpr double kEps = 1.0e-8, kFloor = 1.0e-8;

template <class T> T tmin(T a, T b) { return (val(b) < val(a)) ? b : a; }
template <class T> T tmax(T a, T b) { return (val(a) < val(b)) ? b : a; }
template <class T> T pos(T h) { return val(h) > 0.0 ? h : T(0.0); }
template <class T> T safe_vel(T q, T h) { return val(h) > kEps ? q / h : T(0.0); }
template <class T> T minmod(T a, T b) {
    if (val(a) * val(b) <= 0.0) return T(0.0);
    // (a > 0 ? 1 : -1) * std::min(|a|, |b|): the selected argument carries the derivative
    const bool take_b = dabs(val(b)) < dabs(val(a));
    const T m = take_b ? dabs(b) : dabs(a);
    return val(a) > 0.0 ? m : -m;
}

struct Grid {
    int nx, ny, ng, s;
    double dx, dy, g;
    int bc;
};

template <class T>
void hll(const Grid &gr, T hL, T nL, T tL, T hR, T nR, T tR, T &fh, T &fn, T &ft) {
    const double g = gr.g;
    const T unL = safe_vel(nL, hL), utL = safe_vel(tL, hL);
    const T unR = safe_vel(nR, hR), utR = safe_vel(tR, hR);
    const T cL = dsqrt(g * hL), cR = dsqrt(g * hR);
    const T sL = tmin(unL - cL, unR - cR);
    const T sR = tmax(unL + cL, unR + cR);
    const T FL_h = nL, FL_n = nL * unL + 0.5 * g * hL * hL, FL_t = nL * utL;
    const T FR_h = nR, FR_n = nR * unR + 0.5 * g * hR * hR, FR_t = nR * utR;
    if (val(sL) >= 0.0) {
        fh = FL_h; fn = FL_n; ft = FL_t;
    } else if (val(sR) <= 0.0) {
        fh = FR_h; fn = FR_n; ft = FR_t;
    } else {
        const T inv = T(1.0) / (sR - sL);
        fh = (sR * FL_h - sL * FR_h + sL * sR * (hR - hL)) * inv;
        fn = (sR * FL_n - sL * FR_n + sL * sR * (nR - nL)) * inv;
        ft = (sR * FL_t - sL * FR_t + sL * sR * (tR - tL)) * inv;
    }
}

// Face between cells at k and k+step (normal index), o[0..5] = fm(h, hu, hv), fp(h, hu, hv).
template <class T>
void face(const Grid &gr, bool x, const T *h, const T *hu, const T *hv, const T *B, long k,
          long st, T *o) {
    const double hg = 0.5 * gr.g;
    const long km = k - st, kr = k + st, krr = kr + st;
    const T e0 = h[km] + B[km], e1 = h[k] + B[k], e2 = h[kr] + B[kr], e3 = h[krr] + B[krr];
    const T em = e1 + 0.5 * minmod(e1 - e0, e2 - e1);
    const T ep = e2 - 0.5 * minmod(e2 - e1, e3 - e2);
    const T hum = hu[k] + 0.5 * minmod(hu[k] - hu[km], hu[kr] - hu[k]);
    const T hvm = hv[k] + 0.5 * minmod(hv[k] - hv[km], hv[kr] - hv[k]);
    const T hup = hu[kr] - 0.5 * minmod(hu[kr] - hu[k], hu[krr] - hu[kr]);
    const T hvp = hv[kr] - 0.5 * minmod(hv[kr] - hv[k], hv[krr] - hv[kr]);
    const T bm = B[k], bp = B[kr], bs = tmax(bm, bp);
    const T h_m = pos(em - bm), h_p = pos(ep - bp);
    const T u_m = safe_vel(hum, h_m), v_m = safe_vel(hvm, h_m);
    const T u_p = safe_vel(hup, h_p), v_p = safe_vel(hvp, h_p);
    const T hs_m = pos(em - bs), hs_p = pos(ep - bs);
    T fh, fn, ft;
    if (x)
        hll(gr, hs_m, hs_m * u_m, hs_m * v_m, hs_p, hs_p * u_p, hs_p * v_p, fh, fn, ft);
    else
        hll(gr, hs_m, hs_m * v_m, hs_m * u_m, hs_p, hs_p * v_p, hs_p * u_p, fh, fn, ft);
    const T cm = hg * (h_m * h_m - hs_m * hs_m), cp = hg * (h_p * h_p - hs_p * hs_p);
    o[0] = fh;
    o[3] = fh;
    if (x) {  // hll_x returns (h, hu, hv) with hu normal
        o[1] = fn + cm; o[2] = ft; o[4] = fn + cp; o[5] = ft;
    } else {  // hll_y: hv normal
        o[1] = ft; o[2] = fn + cm; o[4] = ft; o[5] = fn + cp;
    }
}

template <class T>
void ghost(const Grid &gr, T *h, T *hu, T *hv) {
    const int nG = gr.ng, Nx = gr.nx, Ny = gr.ny, s = gr.s, nxt = gr.nx + 2 * gr.ng;
    const bool refl = gr.bc == 0;
    for (int g = 0; g < nG; ++g) {
        const int iL = nG - 1 - g, iR = nG + Nx + g;
        const int sL = refl ? nG + g : nG, sR = refl ? nG + Nx - 1 - g : nG + Nx - 1;
        for (int j = nG; j < nG + Ny; ++j) {
            const long kL = (long)iL * s + j, kSL = (long)sL * s + j;
            const long kR = (long)iR * s + j, kSR = (long)sR * s + j;
            h[kL] = h[kSL]; h[kR] = h[kSR];
            if (hu) {
                hu[kL] = refl ? -hu[kSL] : hu[kSL];
                hu[kR] = refl ? -hu[kSR] : hu[kSR];
                hv[kL] = hv[kSL]; hv[kR] = hv[kSR];
            }
        }
    }
    for (int g = 0; g < nG; ++g) {
        const int jB = nG - 1 - g, jT = nG + Ny + g;
        const int sB = refl ? nG + g : nG, sT = refl ? nG + Ny - 1 - g : nG + Ny - 1;
        for (int i = 0; i < nxt; ++i) {
            const long kB = (long)i * s + jB, kSB = (long)i * s + sB;
            const long kT = (long)i * s + jT, kST = (long)i * s + sT;
            h[kB] = h[kSB]; h[kT] = h[kST];
            if (hu) {
                hu[kB] = hu[kSB]; hu[kT] = hu[kST];
                hv[kB] = refl ? -hv[kSB] : hv[kSB];
                hv[kT] = refl ? -hv[kST] : hv[kST];
            }
        }
    }
}

template <class T>
struct Solver {
    Grid gr;
    long n;
    std::vector<T> B, h1, hu1, hv1, h2, hu2, hv2, r0, r1, r2, f[12];
    Solver(const Grid &g, const T *b) : gr(g), n((long)(g.nx + 2 * g.ng) * (g.ny + 2 * g.ng)) {
        B.assign(b, b + n);
        for (auto *v : {&h1, &hu1, &hv1, &h2, &hu2, &hv2, &r0, &r1, &r2}) v->assign(n, T(0.0));
        for (auto &a : f) a.assign(n, T(0.0));
    }
    void rhs(const T *h, const T *hu, const T *hv) {
        const int nG = gr.ng, s = gr.s;
#pragma omp parallel for schedule(static)
        for (int i = nG - 1; i < nG + gr.nx; ++i)
            for (int j = nG; j < nG + gr.ny; ++j) {
                T o[6];
                const long k = (long)i * s + j;
                face(gr, true, h, hu, hv, B.data(), k, s, o);
                for (int c = 0; c < 6; ++c) f[c][k] = o[c];
            }
#pragma omp parallel for schedule(static)
        for (int i = nG; i < nG + gr.nx; ++i)
            for (int j = nG - 1; j < nG + gr.ny; ++j) {
                T o[6];
                const long k = (long)i * s + j;
                face(gr, false, h, hu, hv, B.data(), k, 1, o);
                for (int c = 0; c < 6; ++c) f[6 + c][k] = o[c];
            }
        const double idx = 1.0 / gr.dx, idy = 1.0 / gr.dy;
        T *r[3] = {r0.data(), r1.data(), r2.data()};
#pragma omp parallel for schedule(static)
        for (int i = nG; i < nG + gr.nx; ++i)
            for (int j = nG; j < nG + gr.ny; ++j) {
                const long k = (long)i * s + j;
                for (int c = 0; c < 3; ++c)
                    r[c][k] = (-idx) * (f[c][k] - f[3 + c][k - s]) - idy * (f[6 + c][k] - f[9 + c][k - 1]);
            }
    }
    bool floor_(T *h, T *hu, T *hv) {
        bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
        for (int i = gr.ng; i < gr.ng + gr.nx; ++i)
            for (int j = gr.ng; j < gr.ng + gr.ny; ++j) {
                const long k = (long)i * gr.s + j;
                if (val(h[k]) < kFloor) { h[k] = T(kFloor); hu[k] = T(0.0); hv[k] = T(0.0); }
                // an output equal to the floor state takes the floored branch's derivative
                // (0) even when it was not floored: the one-sided derivative from below
                if (val(h[k]) == kFloor && val(hu[k]) == 0.0 && val(hv[k]) == 0.0) {
                    h[k] = T(kFloor); hu[k] = T(0.0); hv[k] = T(0.0);
                }
                if (!std::isfinite(val(h[k])) || !std::isfinite(val(hu[k])) || !std::isfinite(val(hv[k])))
                    bad = true;
            }
        return bad;
    }
    bool step(T *h, T *hu, T *hv, double dt) {
        bool bad = false;
        ghost(gr, h, hu, hv);
        rhs(h, hu, hv);
#pragma omp parallel for schedule(static)
        for (long k = 0; k < n; ++k) {
            h1[k] = h[k] + dt * r0[k]; hu1[k] = hu[k] + dt * r1[k]; hv1[k] = hv[k] + dt * r2[k];
        }
        bad |= floor_(h1.data(), hu1.data(), hv1.data());
        ghost(gr, h1.data(), hu1.data(), hv1.data());
        rhs(h1.data(), hu1.data(), hv1.data());
#pragma omp parallel for schedule(static)
        for (long k = 0; k < n; ++k) {
            h2[k] = 0.75 * h[k] + 0.25 * (h1[k] + dt * r0[k]);
            hu2[k] = 0.75 * hu[k] + 0.25 * (hu1[k] + dt * r1[k]);
            hv2[k] = 0.75 * hv[k] + 0.25 * (hv1[k] + dt * r2[k]);
        }
        bad |= floor_(h2.data(), hu2.data(), hv2.data());
        ghost(gr, h2.data(), hu2.data(), hv2.data());
        rhs(h2.data(), hu2.data(), hv2.data());
#pragma omp parallel for schedule(static)
        for (long k = 0; k < n; ++k) {
            h[k] = (1.0 / 3.0) * h[k] + (2.0 / 3.0) * (h2[k] + dt * r0[k]);
            hu[k] = (1.0 / 3.0) * hu[k] + (2.0 / 3.0) * (hu2[k] + dt * r1[k]);
            hv[k] = (1.0 / 3.0) * hv[k] + (2.0 / 3.0) * (hv2[k] + dt * r2[k]);
        }
        bad |= floor_(h, hu, hv);
        ghost(gr, h, hu, hv);
        return bad;
    }
};

}  // namespace

extern "C" {

// Primal run (T = double) with a given dt sequence; arrays in place.  Returns 1 on a
// non-finite state.  Bitwise the oracle's fixed-dt run with the same dts.
int tangent_primal(int nx, int ny, int ng, double dx, double dy, double g, int bc, double *h,
                   double *hu, double *hv, const double *B, const double *dts, long nsteps) {
    Grid gr{nx, ny, ng, ny + 2 * ng, dx, dy, g, bc};
    Solver<double> S(gr, B);
    for (long s = 0; s < nsteps; ++s)
        if (S.step(h, hu, hv, dts[s])) return 1;
    return 0;
}

// Directional derivative of the final state along (dh, dhu, dhv, dB) for a frozen dt
// sequence: on return h/hu/hv hold the primal final state and dh/dhu/dhv its tangent.
int tangent_run(int nx, int ny, int ng, double dx, double dy, double g, int bc, double *h,
                double *hu, double *hv, const double *B, double *dh, double *dhu, double *dhv,
                const double *dB, const double *dts, long nsteps) {
    Grid gr{nx, ny, ng, ny + 2 * ng, dx, dy, g, bc};
    const long n = (long)(nx + 2 * ng) * (ny + 2 * ng);
    std::vector<Dual> Bd(n), H(n), HU(n), HV(n);
    for (long k = 0; k < n; ++k) {
        Bd[k] = Dual(B[k], dB[k]);
        H[k] = Dual(h[k], dh[k]);
        HU[k] = Dual(hu[k], dhu[k]);
        HV[k] = Dual(hv[k], dhv[k]);
    }
    // bathymetry ghosts follow the interior (scalar fill), as at solver setup
    ghost<Dual>(gr, Bd.data(), nullptr, nullptr);
    Solver<Dual> S(gr, Bd.data());
    int rc = 0;
    for (long s = 0; s < nsteps && !rc; ++s) rc = S.step(H.data(), HU.data(), HV.data(), dts[s]);
    for (long k = 0; k < n; ++k) {
        h[k] = H[k].v; hu[k] = HU[k].v; hv[k] = HV[k].v;
        dh[k] = H[k].d; dhu[k] = HU[k].d; dhv[k] = HV[k].d;
    }
    return rc;
}

}  // extern "C"
