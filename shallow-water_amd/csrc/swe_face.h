// swe_face.h -- per-face arithmetic of the MUSCL + hydrostatic reconstruction + HLL scheme,
// written for the FP64 vector ALUs of gfx950.
//
// Every expression keeps the operand order of the reference's optimized kernels
// (include/backend/openmp/omp_kernels.hpp) so that, compiled with -ffp-contract=off, the
// results are bitwise equal to the CPU solvers.  Cited per function.
//
// The face routine is direction-agnostic: "n" is the normal momentum (hu on x faces, hv on
// y faces) and "t" the tangential one.  hll_x (omp_kernels.hpp:50-86) and hll_y (:88-124)
// are exact mirror images under that renaming, so one routine serves both.
#pragma once

#include <hip/hip_runtime.h>

#include "swe_lanes.h"

namespace swe {

constexpr double kEps = 1.0e-8;    // include/constants.hpp:5 (safe_velocity threshold)
constexpr double kFloor = 1.0e-8;  // omp_kernels.hpp:310 (enforce_positivity default floor)

// std::min / std::max, including their NaN and signed-zero behaviour.
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// omp_kernels.hpp:37-39.  v_max_f64 orders -0 below +0 and drops a NaN operand, so
// fmax(h, 0) is +0 for h = -0, 0 and NaN exactly like `h > 0 ? h : 0`.
__device__ __forceinline__ double positive_height(double h) { return fmax(h, 0.0); }

// omp_kernels.hpp:41-43
__device__ __forceinline__ double safe_velocity(double q, double h) {
    return (h > kEps) ? (q / h) : 0.0;
}

// omp_kernels.hpp:45-48.  When a*b > 0, a is non-zero, so (a > 0 ? 1 : -1) * m equals
// copysign(m, a); std::min of two non-negative values equals fmin.  The a*b <= 0 test is
// kept as a product because it can underflow to 0.
__device__ __forceinline__ double minmod(double a, double b) {
    const double r = copysign(fmin(fabs(a), fabs(b)), a);
    return (a * b <= 0.0) ? 0.0 : r;
}

// 0.5 * minmod(a, b) for finite a, b whose product does not underflow unless one of them is
// zero.  Then minmod is the median of (a, b, 0): the smaller magnitude when the signs agree
// and a zero otherwise, of either sign; fma(0.5, m, +0) halves exactly and turns a -0 into
// the +0 that minmod returns.  Five instructions against the seven of 0.5 * minmod.  Under
// the stage kernel's range guard every cell value is 0 or a multiple of 2^-292 of
// magnitude <= 2^190, so every difference of two (eta, hu, hv) values is 0 or at least
// 2^-292 in magnitude and the products stay far from underflow and overflow.
// The min/max are the plain instructions: LLVM's fmin/fmax in IEEE mode first quiet a
// signalling NaN on every operand it cannot prove canonical (DPP moves and loop-carried
// values here), one extra v_max_f64 each; the operands are finite under the guard.  The
// results feed only the fma below (never a DPP move, so no VALU-to-DPP hazard is hidden
// from the compiler).
__device__ __forceinline__ double vmin_f64(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double vmax_f64(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double half_minmod_split(double a, double b) {
    const double m = vmax_f64(vmin_f64(a, b), vmin_f64(vmax_f64(a, b), 0.0));
    return fma(0.5, m, 0.0);
}

// Result of one face: mass flux, tangential-momentum flux, and the normal-momentum flux as
// seen from the minus (left/bottom) and plus (right/top) cell, which differ by the
// hydrostatic pressure correction (omp_kernels.hpp:179-189, :247-257).
struct FaceFlux {
    double fh, ft, fn_m, fn_p;
};

// 1/d from the hardware reciprocal estimate refined by two Newton steps (relative error
// within about one ulp).  Used only by the FAST precision mode.
__device__ __forceinline__ double recip(double d) {
    const double r0 = __builtin_amdgcn_rcp(d);
    const double r1 = fma(r0, fma(-d, r0, 1.0), r0);
    return fma(r1, fma(-d, r1, 1.0), r1);
}

// ---- IEEE division and square root, split for operands in a guarded range -------------
//
// On gfx950 an f64 division n/d is the sequence
//   ds = div_scale(d, d, n); r = rcp(ds); ns = div_scale(n, d, n)
//   e = fma(-ds, r, 1); r = fma(r, e, r); e = fma(-ds, r, 1); r = fma(r, e, r)
//   q = ns * r; rem = fma(-ds, q, ns); q = div_fmas(rem, r, q); n/d = div_fixup(q, d, n)
// and sqrt(x) is rsq + six fma/mul Newton steps, bracketed by a rescale for x < 2^-767 and
// a select for x in {+-0, +inf}.  When n and d are normal, d > 0, n/d is normal and the
// exponents are far from the limits (|n| >= 2^-969, exponent(n) - exponent(d) < 768),
// div_scale returns its operand unchanged with VCC = 0, div_fmas is a plain fma and
// div_fixup only applies the sign rule sign(n) ^ sign(d); sqrt's rescale is the identity
// for x >= 2^-767.  The split forms below compute exactly those reduced sequences, so the
// results are the same bits as `/` and sqrt() -- and the denominator part (rcp + 4 fma) is
// shared by the two divisions that have the same depth as denominator.  Callers use them
// only under the range guard of the stage kernel (every cell value 0 or of magnitude in
// [2^-240, 2^190]), which bounds every numerator, denominator and sqrt argument of a face
// well inside those conditions.
__device__ __forceinline__ double div_denominator(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// n / d given r = div_denominator(d), for d > 0.  The reduced sequence is
// fma(rem, r, q) with rem = fma(-d, q, n), q = n*r, and the sign of n; here the remainder
// is taken negated, -fma(d, q, -n): round-to-nearest is symmetric, so a non-zero remainder
// is the same value, and an exact zero remainder comes out as -0 instead of +0.  That makes
// the last fma return q's sign for a zero result (q = n*r with r > 0 carries the sign of n,
// -0 included), so the sign rule needs no copysign; for a non-zero result a zero term adds
// nothing whatever its sign.
__device__ __forceinline__ double div_numerator(double n, double d, double r) {
    const double q = n * r;
    return fma(-fma(d, q, -n), r, q);
}

// sqrt(x) for x >= 2^-767 (finite, non-zero).
__device__ __forceinline__ double sqrt_positive(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

// sqrt(x) for x = 0 or x >= 2^-767 (finite).
__device__ __forceinline__ double sqrt_unscaled(double x) {
    const double g = sqrt_positive(x);
    return x == 0.0 ? x : g;
}

// The range of cell values under which the split forms are exact (see above): zero, or a
// magnitude in [2^-240, 2^190].  NaN, infinities and denormals fail it.
__device__ __forceinline__ bool split_ok(double v) {
    const double a = fabs(v);
    return (a >= 0x1p-240 && a <= 0x1p190) || v == 0.0;
}

// The lanes of the wave where split_ok(v) holds, one ballot per compare (the masks combine
// on the scalar unit instead of as 0/1 vector values).
__device__ __forceinline__ unsigned long long split_ok_lanes(double v) {
    const double a = fabs(v);
    return (__builtin_amdgcn_ballot_w64(a >= 0x1p-240) & __builtin_amdgcn_ballot_w64(a <= 0x1p190)) |
           __builtin_amdgcn_ballot_w64(v == 0.0);
}

// Division and square root in the exact mode: the split forms when SPLIT (range proven),
// the compiler's full sequences otherwise.
template <bool SPLIT>
__device__ __forceinline__ double ieee_sqrt(double x) {
    if constexpr (SPLIT) return sqrt_unscaled(x);
    return sqrt(x);
}

// ---- exact mode, split forms (range guard holds) ---------------------------------------
//
// The same arithmetic as hll/face_flux below with SPLIT = true, organised so that
// wave-uniform facts skip work that cannot change a result:
//  * WET (every lane's two depths > eps): safe_velocity's selects are the identity and the
//    sqrt arguments g*h are positive, so the zero test of the square root is dropped;
//  * a flat bottom across the face in every lane (b- == b+): b* equals both bottoms, so the
//    hydrostatic depths are the face depths themselves (positive_height maps a -0/+0
//    difference to +0 either way), HLL reuses their reciprocals, and the pressure
//    correction half_g*(h*h - hs*hs) is exactly +0 (fn + corr == fn + 0.0).
// Each variant is a separate body, and the empty volatile asm statements around the WET
// bodies keep the compiler from merging the two bodies back into one with selects (which
// would cost more than the branch saves and raises register pressure).

// HLL given the reciprocals rL = div_denominator(hL), rR = div_denominator(hR).
template <bool WET>
__device__ __forceinline__ void hll_split(double half_g, double g, double hL, double nL,
                                          double tL, double hR, double nR, double tR, double rL,
                                          double rR, double &fh, double &fn, double &ft,
                                          unsigned long long skip) {
    double unL = div_numerator(nL, hL, rL), utL = div_numerator(tL, hL, rL);
    double unR = div_numerator(nR, hR, rR), utR = div_numerator(tR, hR, rR);
    double cL, cR;
    if constexpr (WET) {
        cL = sqrt_positive(g * hL);
        cR = sqrt_positive(g * hR);
    } else {
        unL = hL > kEps ? unL : 0.0;
        utL = hL > kEps ? utL : 0.0;
        unR = hR > kEps ? unR : 0.0;
        utR = hR > kEps ? utR : 0.0;
        cL = sqrt_unscaled(g * hL);
        cR = sqrt_unscaled(g * hR);
    }
    const double sL = fmin(unL - cL, unR - cR);
    const double sR = fmax(unL + cL, unR + cR);
    const double FL_h = nL;
    const double FL_n = nL * unL + half_g * hL * hL;
    const double FL_t = nL * utL;
    const double FR_h = nR;
    const double FR_n = nR * unR + half_g * hR * hR;
    const double FR_t = nR * utR;
    auto star = [&](double &h_, double &n_, double &t_) {
        const double d = sR - sL;
        const double inv = div_numerator(1.0, d, div_denominator(d));
        h_ = (sR * FL_h - sL * FR_h + sL * sR * (hR - hL)) * inv;
        n_ = (sR * FL_n - sL * FR_n + sL * sR * (nR - nL)) * inv;
        t_ = (sR * FL_t - sL * FR_t + sL * sR * (tR - tL)) * inv;
    };
    if (wave_all2(sL < 0.0, sR > 0.0, skip)) {
        star(fh, fn, ft);
    } else if (sL >= 0.0) {
        fh = FL_h;
        fn = FL_n;
        ft = FL_t;
    } else if (sR <= 0.0) {
        fh = FR_h;
        fn = FR_n;
        ft = FR_t;
    } else {
        star(fh, fn, ft);
    }
}

// Face with b- == b+ in every lane; WET: every lane's face depths > eps.
template <bool WET>
__device__ __forceinline__ FaceFlux face_flat_split(double half_g, double g, double h_m,
                                                    double n_m, double t_m, double h_p,
                                                    double n_p, double t_p,
                                                    unsigned long long skip) {
    const double rm = div_denominator(h_m), rp = div_denominator(h_p);
    double un_m = div_numerator(n_m, h_m, rm), ut_m = div_numerator(t_m, h_m, rm);
    double un_p = div_numerator(n_p, h_p, rp), ut_p = div_numerator(t_p, h_p, rp);
    if constexpr (!WET) {
        un_m = h_m > kEps ? un_m : 0.0;
        ut_m = h_m > kEps ? ut_m : 0.0;
        un_p = h_p > kEps ? un_p : 0.0;
        ut_p = h_p > kEps ? ut_p : 0.0;
    }
    double fh, fn, ft;
    hll_split<WET>(half_g, g, h_m, h_m * un_m, h_m * ut_m, h_p, h_p * un_p, h_p * ut_p, rm, rp,
                   fh, fn, ft, skip);
    FaceFlux f;
    f.fh = fh;
    f.ft = ft;
    f.fn_m = fn + 0.0;
    f.fn_p = fn + 0.0;
    return f;
}

// Face with a step in the bottom; WET: every lane's hydrostatic depths hs > eps, hence
// also its face depths (hs = max(eta - b*, 0) <= max(eta - b, 0) = h since b* >= b).
template <bool WET>
__device__ __forceinline__ FaceFlux face_step_split(double half_g, double g, double h_m,
                                                    double n_m, double t_m, double hs_m,
                                                    double h_p, double n_p, double t_p,
                                                    double hs_p, unsigned long long skip) {
    const double rm = div_denominator(h_m), rp = div_denominator(h_p);
    double un_m = div_numerator(n_m, h_m, rm), ut_m = div_numerator(t_m, h_m, rm);
    double un_p = div_numerator(n_p, h_p, rp), ut_p = div_numerator(t_p, h_p, rp);
    if constexpr (!WET) {
        un_m = h_m > kEps ? un_m : 0.0;
        ut_m = h_m > kEps ? ut_m : 0.0;
        un_p = h_p > kEps ? un_p : 0.0;
        ut_p = h_p > kEps ? ut_p : 0.0;
    }
    double fh, fn, ft;
    hll_split<WET>(half_g, g, hs_m, hs_m * un_m, hs_m * ut_m, hs_p, hs_p * un_p, hs_p * ut_p,
                   div_denominator(hs_m), div_denominator(hs_p), fh, fn, ft, skip);
    const double corr_m = half_g * (h_m * h_m - hs_m * hs_m);
    const double corr_p = half_g * (h_p * h_p - hs_p * hs_p);
    FaceFlux f;
    f.fh = fh;
    f.ft = ft;
    f.fn_m = fn + corr_m;
    f.fn_p = fn + corr_p;
    return f;
}

// face_flux<false, true> (below): the same values.
__device__ __forceinline__ FaceFlux face_flux_split(double half_g, double g, double eta_m,
                                                   double n_m, double t_m, double b_m,
                                                   double eta_p, double n_p, double t_p,
                                                   double b_p, unsigned long long skip) {
    const double h_m = positive_height(eta_m - b_m);
    const double h_p = positive_height(eta_p - b_p);
    if (wave_all(b_m == b_p, skip)) {
        if (wave_all2(h_m > kEps, h_p > kEps, skip)) {
            asm volatile("");
            const FaceFlux f =
                face_flat_split<true>(half_g, g, h_m, n_m, t_m, h_p, n_p, t_p, skip);
            asm volatile("");
            return f;
        }
        return face_flat_split<false>(half_g, g, h_m, n_m, t_m, h_p, n_p, t_p, skip);
    }
    const double b_star = fmax(b_m, b_p);
    const double hs_m = positive_height(eta_m - b_star);
    const double hs_p = positive_height(eta_p - b_star);
    if (wave_all2(hs_m > kEps, hs_p > kEps, skip)) {
        asm volatile("");
        const FaceFlux f =
            face_step_split<true>(half_g, g, h_m, n_m, t_m, hs_m, h_p, n_p, t_p, hs_p, skip);
        asm volatile("");
        return f;
    }
    return face_step_split<false>(half_g, g, h_m, n_m, t_m, hs_m, h_p, n_p, t_p, hs_p, skip);
}

// HLL on the hydrostatically reconstructed states (omp_kernels.hpp:50-86 / :88-124).
//
// FAST = false: the reference arithmetic, every division IEEE (bitwise parity).
// FAST = true : the velocities of the reconstructed states are taken from the caller
//   (vnL.. = the face velocities before the hydrostatic depth change, so hs*u/hs -> u, a
//   1-ulp difference) and 1/(sR - sL) is a refined reciprocal.  Results then differ from
//   the reference by a few ulps per step, which the tests bound after N steps.
template <bool FAST, bool SPLIT>
__device__ __forceinline__ void hll(double half_g, double g, double hL, double nL, double tL,
                                    double hR, double nR, double tR, double vnL, double vtL,
                                    double vnR, double vtR, double &fh, double &fn,
                                    double &ft, unsigned long long skip) {
    double unL, utL, unR, utR;
    if constexpr (FAST) {
        unL = hL > kEps ? vnL : 0.0;
        utL = hL > kEps ? vtL : 0.0;
        unR = hR > kEps ? vnR : 0.0;
        utR = hR > kEps ? vtR : 0.0;
    } else if constexpr (SPLIT) {
        const double rL = div_denominator(hL), rR = div_denominator(hR);
        const double qnL = div_numerator(nL, hL, rL), qtL = div_numerator(tL, hL, rL);
        const double qnR = div_numerator(nR, hR, rR), qtR = div_numerator(tR, hR, rR);
        unL = hL > kEps ? qnL : 0.0;
        utL = hL > kEps ? qtL : 0.0;
        unR = hR > kEps ? qnR : 0.0;
        utR = hR > kEps ? qtR : 0.0;
    } else if (wave_all2(hL > kEps, hR > kEps, skip)) {
        // whole wave wet: safe_velocity takes its division branch in every lane
        unL = nL / hL;
        utL = tL / hL;
        unR = nR / hR;
        utR = tR / hR;
    } else {
        unL = safe_velocity(nL, hL);
        utL = safe_velocity(tL, hL);
        unR = safe_velocity(nR, hR);
        utR = safe_velocity(tR, hR);
    }

    const double cL = ieee_sqrt<SPLIT>(g * hL);
    const double cR = ieee_sqrt<SPLIT>(g * hR);

    // fmin/fmax instead of std::min/max: they differ only in the sign of a zero result or
    // on NaN, and a zero sL (sR) takes the sL >= 0 (sR <= 0) branch where its value is unused.
    const double sL = fmin(unL - cL, unR - cR);
    const double sR = fmax(unL + cL, unR + cR);

    const double FL_h = nL;
    const double FL_n = nL * unL + half_g * hL * hL;
    const double FL_t = nL * utL;
    const double FR_h = nR;
    const double FR_n = nR * unR + half_g * hR * hR;
    const double FR_t = nR * utR;

    // Star (subcritical) branch; when every lane of the wave is subcritical the selects of
    // the one-sided branches are skipped (same values, fewer instructions).
    auto star = [&](double &h_, double &n_, double &t_) {
        double inv;
        if constexpr (FAST) {
            inv = recip(sR - sL);
        } else if constexpr (SPLIT) {
            const double d = sR - sL;
            inv = div_numerator(1.0, d, div_denominator(d));
        } else {
            inv = 1.0 / (sR - sL);
        }
        h_ = (sR * FL_h - sL * FR_h + sL * sR * (hR - hL)) * inv;
        n_ = (sR * FL_n - sL * FR_n + sL * sR * (nR - nL)) * inv;
        t_ = (sR * FL_t - sL * FR_t + sL * sR * (tR - tL)) * inv;
    };
    if (wave_all2(sL < 0.0, sR > 0.0, skip)) {
        star(fh, fn, ft);
    } else if (sL >= 0.0) {
        fh = FL_h;
        fn = FL_n;
        ft = FL_t;
    } else if (sR <= 0.0) {
        fh = FR_h;
        fn = FR_n;
        ft = FR_t;
    } else {
        star(fh, fn, ft);
    }
}

// One face from the MUSCL-reconstructed minus/plus values of (eta, n, t) and the cell
// bottoms on either side (omp_kernels.hpp:154-189 for x, :222-257 for y).
// skip: lanes whose inputs are not meaningful (their results are never used) and which
// must not take part in the wave-uniform branch decisions.
template <bool FAST, bool SPLIT>
__device__ __forceinline__ FaceFlux face_flux(double half_g, double g, double eta_m, double n_m,
                                              double t_m, double b_m, double eta_p, double n_p,
                                              double t_p, double b_p,
                                              unsigned long long skip = 0ull) {
    if constexpr (SPLIT && !FAST)
        return face_flux_split(half_g, g, eta_m, n_m, t_m, b_m, eta_p, n_p, t_p, b_p, skip);
    // fmax: differs from std::max only for (b_m, b_p) = (-0, +0), where eta - b_star then
    // differs only in the sign of a zero difference, which positive_height maps to +0.
    const double b_star = fmax(b_m, b_p);

    const double h_m = positive_height(eta_m - b_m);
    const double h_p = positive_height(eta_p - b_p);

    double un_m, ut_m, un_p, ut_p;
    if constexpr (FAST) {
        // one refined reciprocal per depth instead of two divisions
        const double rm = recip(h_m), rp = recip(h_p);
        un_m = h_m > kEps ? n_m * rm : 0.0;
        ut_m = h_m > kEps ? t_m * rm : 0.0;
        un_p = h_p > kEps ? n_p * rp : 0.0;
        ut_p = h_p > kEps ? t_p * rp : 0.0;
    } else if constexpr (SPLIT) {
        const double rm = div_denominator(h_m), rp = div_denominator(h_p);
        const double qnm = div_numerator(n_m, h_m, rm), qtm = div_numerator(t_m, h_m, rm);
        const double qnp = div_numerator(n_p, h_p, rp), qtp = div_numerator(t_p, h_p, rp);
        un_m = h_m > kEps ? qnm : 0.0;
        ut_m = h_m > kEps ? qtm : 0.0;
        un_p = h_p > kEps ? qnp : 0.0;
        ut_p = h_p > kEps ? qtp : 0.0;
    } else if (wave_all2(h_m > kEps, h_p > kEps, skip)) {
        // whole wave wet: safe_velocity takes its division branch in every lane
        un_m = n_m / h_m;
        ut_m = t_m / h_m;
        un_p = n_p / h_p;
        ut_p = t_p / h_p;
    } else {
        un_m = safe_velocity(n_m, h_m);
        ut_m = safe_velocity(t_m, h_m);
        un_p = safe_velocity(n_p, h_p);
        ut_p = safe_velocity(t_p, h_p);
    }

    const double hs_m = positive_height(eta_m - b_star);
    const double hs_p = positive_height(eta_p - b_star);

    double fh, fn, ft;
    hll<FAST, SPLIT>(half_g, g, hs_m, hs_m * un_m, hs_m * ut_m, hs_p, hs_p * un_p, hs_p * ut_p,
                     un_m, ut_m, un_p, ut_p, fh, fn, ft, skip);

    const double corr_m = half_g * (h_m * h_m - hs_m * hs_m);
    const double corr_p = half_g * (h_p * h_p - hs_p * hs_p);

    FaceFlux f;
    f.fh = fh;
    f.ft = ft;
    f.fn_m = fn + corr_m;
    f.fn_p = fn + corr_p;
    return f;
}

// Divergence of one component (omp_kernels.hpp:281-283):
//   -inv_dx*(fxm[k] - fxp[k-s]) - inv_dy*(fym[k] - fyp[k-1])
__device__ __forceinline__ double divergence(double inv_dx, double inv_dy, double fxm,
                                             double fxp_left, double fym, double fyp_below) {
    return -inv_dx * (fxm - fxp_left) - inv_dy * (fym - fyp_below);
}

// SSP-RK3 stage combinations (omp_solver.hpp:264-266, :290-292, :316-320).  `in` is the
// stage input at the cell (U, U1 or U2), `u0` is U^n.
template <int STAGE>
__device__ __forceinline__ double rk_combine(double u0, double in, double dt, double rhs) {
    if constexpr (STAGE == 1) {
        return in + dt * rhs;
    } else if constexpr (STAGE == 2) {
        return 0.75 * u0 + 0.25 * (in + dt * rhs);
    } else {
        return (1.0 / 3.0) * u0 + (2.0 / 3.0) * (in + dt * rhs);
    }
}

// Per-cell wave speed of compute_stable_dt (omp_kernels.hpp:297-302): max(|u|+c, |v|+c).
__device__ __forceinline__ double cell_speed(double g, double h, double hu, double hv,
                                             double running) {
    const double u = safe_velocity(hu, h);
    const double v = safe_velocity(hv, h);
    const double c = sqrt(g * smax(0.0, h));
    // all operands are >= +0 here, where fmax equals std::max
    running = fmax(running, fabs(u) + c);
    running = fmax(running, fabs(v) + c);
    return running;
}

// compute_stable_dt's final expression (omp_kernels.hpp:305-306).
__device__ __forceinline__ double dt_from_speed(double max_speed, double cfl, double dx,
                                                double dy) {
    if (max_speed <= kEps) return 1.0e100;
    return cfl * smin(dx, dy) / max_speed;
}

}  // namespace swe
