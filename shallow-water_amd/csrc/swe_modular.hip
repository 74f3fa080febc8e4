// swe_modular.hip -- fused SSP-RK3 stage kernels of the reference's modular serial solver
// (SerialSolver<ReflectingWalls, Recon, Riemann, SSPRK3, Bathymetry>,
// include/backend/serial/serial_solver.hpp) for the combinations its SolverAssembly
// dispatches (include/solver_assembly/solver_assembly.hpp:315-500):
//   no bathymetry  : {PiecewiseConst, MUSCL} x {HLL, Rusanov, ROE}  (flux_assembly.hpp)
//   with bathymetry: {PiecewiseConst, MUSCL} x HLL                  (flux_assembly_bathy.hpp)
//
// The work decomposition is the optimized stage kernel's (swe_stage.hip): a wave owns a
// 60-column strip plus a 2-column halo and marches down x rows with the radius-2 x stencil
// in registers (each x face computed once, the march unrolled 4x with the register window
// renamed), y neighbours in adjacent lanes exchanged with DPP moves (each y face computed
// once), ghost cells synthesized by mirrored loads with
// the reflecting walls' sign flips.  The arithmetic is the modular solver's own, operand
// by operand, and differs from the optimized path's:
//   * physical fluxes F (x) and G (y) of riemann.hpp:12-19 (u*u*h for x, h*v*v for y);
//   * velocities of CellState::u/v (h > eps) inside the Riemann solvers, but
//     flux_assembly_bathy.hpp:233-240's `h < eps ? 0 : q/h` for the hydrostatic states;
//   * HLL divides its star flux by (sR - sL) (HLL.hpp:19); Rusanov and ROE as written;
//   * stage combinations of ssp_rk3.hpp: U + dt L, 3/4 U + 1/4 U1 + dt/4 L,
//     1/3 U + 2/3 U2 + (2 dt/3) L;
//   * the positivity floor and non-finite check only with bathymetry (serial_solver.hpp:
//     190-206); the plain path has neither.
// Division and sqrt: the split forms of swe_face.h (same bits as `/` and sqrt()) for waves
// whose cells all pass the modular range guard (mod_in_range / mod_zero_or_in_range below),
// the compiler's IEEE sequences otherwise.  Where the reference's HLL throws (NaN wave speeds,
// HLL.hpp:23), the stage raises the non-finite flag instead.
#include <type_traits>

#include "swe_face.h"
#include "swe_internal.h"
#include "swe_lanes.h"

namespace swe {
namespace {

enum { kReconMUSCL = 0, kReconPC = 1 };
enum { kHLL = 0, kRusanov = 1, kROE = 2 };

// A state as (depth or eta, normal momentum, tangential momentum) of the face direction.
struct NT {
    double h, n, t;
};

// ---- division and square root -----------------------------------------------------------
// SPLIT: the reduced sequences of swe_face.h (rcp + 4 fma per denominator, shared by the
// divisions with that denominator; mul, fma, fma per numerator; sqrt without its rescale).
// They equal `/` and sqrt() bit for bit when d > 0, n/d and n are normal or zero, the
// exponents are far from the limits and sqrt's argument is 0 or >= 2^-767.  The modular
// range guard makes that hold for every division and root these schemes form at a wave
// whose loaded depths are in [2^-100, 2^100] and momenta and bottoms 0 or in that range:
//  * reconstructed values are multiples of 2^-153 of magnitude <= 2^102; minmod-limited
//    MUSCL keeps a reconstructed depth between its neighbours, so plain depths stay
//    >= 2^-153; hydrostatic depths max(eta - b, 0) are 0 or >= 2^-153;
//  * velocities q/h are taken only for h > eps = 1e-8 (the selects discard the others):
//    |u| <= 2^102 / 2^-26.5 < 2^129, nonzero |u| >= 2^-255;
//  * sqrt(g h) with g in [2^-10, 2^10] (checked on the host): 0 or >= 2^-82;
//  * HLL's star divides by d = sR - sL > 0 where sL < 0 < sR, and d >= 2 max(cL, cR), so
//    d >= 2^-81 unless both sides are dry (then sL = sR = 0 and no star); every flux
//    component is < 2^234 and every quotient (N/d, |N/d| <= |FL| + |FR| + |sL dU|) < 2^236,
//    while a nonzero N is a rounded sum of products of at least 2^-82 * 2^-255, far above
//    2^-969 with a quotient above 2^-600;
//  * ROE divides by sqrt(h1) + sqrt(h2) >= 2^-77 and by c = sqrt(g h) >= 2^-81, with
//    numerators below 2^240.
template <bool SPLIT>
__device__ __forceinline__ double mrecip(double d) {
    if constexpr (SPLIT) return div_denominator(d);
    return 0.0;  // unused: the IEEE division needs no separate reciprocal
}
template <bool SPLIT>
__device__ __forceinline__ double mdiv(double n, double d, double r) {
    if constexpr (SPLIT) return div_numerator(n, d, r);
    return n / d;
}

// The lanes whose cell passes the modular range guard (above), one ballot per compare.
__device__ __forceinline__ unsigned long long mod_in_range(double v) {
    const double a = fabs(v);
    return __builtin_amdgcn_ballot_w64(a >= 0x1p-100) & __builtin_amdgcn_ballot_w64(a <= 0x1p100);
}
__device__ __forceinline__ unsigned long long mod_zero_or_in_range(double v) {
    return mod_in_range(v) | __builtin_amdgcn_ballot_w64(v == 0.0);
}

// WET (split hydrostatic faces only): every counted lane of the wave has both hydrostatic
// depths > eps, so the velocity selects below take their division in every lane and the
// square roots' arguments are positive -- the selects and the zero test are dropped (same
// values; measured 2-3.6% faster with bathymetry).
template <bool SPLIT, bool WET>
__device__ __forceinline__ double msqrt(double x) {
    if constexpr (SPLIT && WET) return sqrt_positive(x);
    return ieee_sqrt<SPLIT>(x);
}

// CellState::u / v (cell_state.hpp:19-26); r = mrecip<SPLIT>(h)
template <bool SPLIT, bool WET = false>
__device__ __forceinline__ double cs_vel(double q, double h, double r) {
    if constexpr (WET) return mdiv<SPLIT>(q, h, r);
    return (h > kEps) ? mdiv<SPLIT>(q, h, r) : 0.0;
}
// flux_assembly_bathy.hpp:233-240; r = mrecip<SPLIT>(h)
template <bool SPLIT, bool WET = false>
__device__ __forceinline__ double fa_vel(double q, double h, double r) {
    if constexpr (WET) return mdiv<SPLIT>(q, h, r);
    return (h < kEps) ? 0.0 : mdiv<SPLIT>(q, h, r);
}
// flux_assembly_bathy.hpp:231 std::max(0.0, h)
__device__ __forceinline__ double fa_positive(double h) { return (0.0 < h) ? h : 0.0; }

// riemann.hpp:12-19 in normal/tangential form: x uses F = (hu, u*u*h + g/2 h^2, hu*v);
// y uses G = (hv, hu*v, h*v*v + g/2 h^2), i.e. (n, t*un, h*un*un + ...).
template <int DIR, bool SPLIT, bool WET>
__device__ __forceinline__ NT phys(double half_g, const NT &U, double r) {
    const double un = cs_vel<SPLIT, WET>(U.n, U.h, r), ut = cs_vel<SPLIT, WET>(U.t, U.h, r);
    NT f;
    f.h = U.n;
    if (DIR == 0) {
        f.n = un * un * U.h + half_g * U.h * U.h;
        f.t = U.n * ut;
    } else {
        f.n = U.h * un * un + half_g * U.h * U.h;
        f.t = U.t * un;
    }
    return f;
}

// HLL.hpp:9-45
template <int DIR, bool SPLIT, bool WET>
__device__ __forceinline__ NT hll_mod(double g, double half_g, const NT &L, const NT &R,
                                      int &bad) {
    const double rL = mrecip<SPLIT>(L.h), rR = mrecip<SPLIT>(R.h);
    const double uL = cs_vel<SPLIT, WET>(L.n, L.h, rL), uR = cs_vel<SPLIT, WET>(R.n, R.h, rR);
    const double cL = msqrt<SPLIT, WET>(g * L.h), cR = msqrt<SPLIT, WET>(g * R.h);
    const double sR = smax(uL + cL, uR + cR);
    const double sL = smin(uL - cL, uR - cR);
    const NT FL = phys<DIR, SPLIT, WET>(half_g, L, rL), FR = phys<DIR, SPLIT, WET>(half_g, R, rR);
    if (sL >= 0.0) return FL;
    if (sL < 0.0 && sR > 0.0) {
        const double d = sR - sL, ss = sL * sR;
        const double rd = mrecip<SPLIT>(d);
        return NT{mdiv<SPLIT>(sR * FL.h - sL * FR.h + ss * (R.h - L.h), d, rd),
                  mdiv<SPLIT>(sR * FL.n - sL * FR.n + ss * (R.n - L.n), d, rd),
                  mdiv<SPLIT>(sR * FL.t - sL * FR.t + ss * (R.t - L.t), d, rd)};
    }
    if (sR <= 0.0) return FR;
    bad = 1;  // the reference throws "HLL could not provide dissipation speed."
    return FL;
}

// rusanov.hpp:11-30
template <int DIR, bool SPLIT>
__device__ __forceinline__ NT rusanov(double g, double half_g, const NT &L, const NT &R) {
    const double rL = mrecip<SPLIT>(L.h), rR = mrecip<SPLIT>(R.h);
    const double uL = cs_vel<SPLIT>(L.n, L.h, rL), uR = cs_vel<SPLIT>(R.n, R.h, rR);
    const double hL = (L.h < 0.0) ? 0.0 : L.h, hR = (R.h < 0.0) ? 0.0 : R.h;  // std::max(h, 0)
    const double a = smax(fabs(uL) + ieee_sqrt<SPLIT>(g * hL), fabs(uR) + ieee_sqrt<SPLIT>(g * hR));
    const NT FL = phys<DIR, SPLIT, false>(half_g, L, rL), FR = phys<DIR, SPLIT, false>(half_g, R, rR);
    const double ha = 0.5 * a;
    return NT{0.5 * (FL.h + FR.h) - ha * (R.h - L.h), 0.5 * (FL.n + FR.n) - ha * (R.n - L.n),
              0.5 * (FL.t + FR.t) - ha * (R.t - L.t)};
}

// ROE.hpp:66-69, given q1 = sqrt(h1), q2 = sqrt(h2), r = mrecip<SPLIT>(q1 + q2)
template <bool SPLIT>
__device__ __forceinline__ double roe_avg(double u1, double q1, double u2, double q2, double r) {
    return mdiv<SPLIT>((q1 * u1) + (q2 * u2), q1 + q2, r);
}

// ROE.hpp:9-61.  In normal/tangential form both directions read the same: eigenvectors
// r1 = (1, l1, ut), r2 = (0, 0, 1), r3 = (1, l3, ut); the reference's array operators
// multiply element * scalar and add commutatively, so these sums are its sums.
template <int DIR, bool SPLIT, bool WET>
__device__ __forceinline__ NT roe(double g, double half_g, const NT &L, const NT &R) {
    const double h = (L.h + R.h) / 2.0;
    const double c = msqrt<SPLIT, WET>(g * h);
    const double rL = mrecip<SPLIT>(L.h), rR = mrecip<SPLIT>(R.h);
    const double qL = msqrt<SPLIT, WET>(L.h), qR = msqrt<SPLIT, WET>(R.h);
    const double rq = mrecip<SPLIT>(qL + qR);
    const double un = roe_avg<SPLIT>(cs_vel<SPLIT, WET>(L.n, L.h, rL), qL,
                                     cs_vel<SPLIT, WET>(R.n, R.h, rR), qR, rq);
    const double ut = roe_avg<SPLIT>(cs_vel<SPLIT, WET>(L.t, L.h, rL), qL,
                                     cs_vel<SPLIT, WET>(R.t, R.h, rR), qR, rq);
    const double l1 = un - c, l2 = un, l3 = un + c;
    const double dh = R.h - L.h, dn = R.n - L.n, dt = R.t - L.t;
    const double w = mdiv<SPLIT>(dn - un * dh, c, mrecip<SPLIT>(c));
    const double a1 = 0.5 * (dh - w);
    const double a3 = 0.5 * (dh + w);
    const double a2 = dt - ut * dh;
    const double s1 = fabs(l1) * a1, s2 = fabs(l2) * a2, s3 = fabs(l3) * a3;
    const double sh = ((1.0 * s1) + (0.0 * s2)) + (1.0 * s3);
    const double sn = ((l1 * s1) + (0.0 * s2)) + (l3 * s3);
    const double st_ = ((ut * s1) + (1.0 * s2)) + (ut * s3);
    const NT FL = phys<DIR, SPLIT, WET>(half_g, L, rL), FR = phys<DIR, SPLIT, WET>(half_g, R, rR);
    return NT{0.5 * (FL.h + FR.h) - sh * 0.5, 0.5 * (FL.n + FR.n) - sn * 0.5,
              0.5 * (FL.t + FR.t) - st_ * 0.5};
}

// The selected Riemann solver.
template <int DIR, int RIEMANN, bool SPLIT, bool WET>
__device__ __forceinline__ NT riemann_v(double g, double half_g, const NT &L, const NT &R,
                                        int &bad) {
    if constexpr (RIEMANN == kRusanov) return rusanov<DIR, SPLIT>(g, half_g, L, R);
    else if constexpr (RIEMANN == kROE) return roe<DIR, SPLIT, WET>(g, half_g, L, R);
    else return hll_mod<DIR, SPLIT, WET>(g, half_g, L, R, bad);
}

// One face from the reconstructed minus/plus states (w = h, or eta with bathymetry) and
// the cell bottoms.  Returns (fh, ft, fn_minus, fn_plus) like the optimized face.  skip:
// lanes left out of the wave vote (the y faces of lanes 62 and 63 read zero-filled moves).
template <int DIR, int RIEMANN, bool BATHY, bool SPLIT>
__device__ __forceinline__ FaceFlux face_mod(double g, double half_g, const NT &m, double b_m,
                                             const NT &p, double b_p, int &bad,
                                             unsigned long long skip) {
    FaceFlux f;
    if constexpr (!BATHY) {
        // flux_assembly.hpp:19-25 (the all-wet body of the hydrostatic branch below measured
        // 3.5% slower here with piecewise-constant HLL and within noise otherwise)
        const NT F = riemann_v<DIR, RIEMANN, SPLIT, false>(g, half_g, m, p, bad);
        f.fh = F.h;
        f.ft = F.t;
        f.fn_m = F.n;
        f.fn_p = F.n;
    } else {
        // flux_assembly_bathy.hpp:35-88
        const double b_star = smax(b_m, b_p);
        const double h_m = fa_positive(m.h - b_m), h_p = fa_positive(p.h - b_p);
        const double r_m = mrecip<SPLIT>(h_m), r_p = mrecip<SPLIT>(h_p);
        const double hs_m = fa_positive(m.h - b_star), hs_p = fa_positive(p.h - b_star);
        // WET: hs > eps on both sides, hence also h >= hs > eps for the face velocities
        auto body = [&](auto wet_tag) {
            constexpr bool WET = decltype(wet_tag)::value;
            const double un_m = fa_vel<SPLIT, WET>(m.n, h_m, r_m);
            const double ut_m = fa_vel<SPLIT, WET>(m.t, h_m, r_m);
            const double un_p = fa_vel<SPLIT, WET>(p.n, h_p, r_p);
            const double ut_p = fa_vel<SPLIT, WET>(p.t, h_p, r_p);
            const NT Um{hs_m, hs_m * un_m, hs_m * ut_m}, Up{hs_p, hs_p * un_p, hs_p * ut_p};
            return riemann_v<DIR, RIEMANN, SPLIT, WET>(g, half_g, Um, Up, bad);
        };
        NT F;
        if constexpr (SPLIT && RIEMANN != kRusanov) {
            if (wave_all2(hs_m > kEps, hs_p > kEps, skip)) {
                asm volatile("");
                F = body(std::true_type{});
                asm volatile("");
            } else {
                F = body(std::false_type{});
            }
        } else {
            F = body(std::false_type{});
        }
        f.fh = F.h;
        f.ft = F.t;
        f.fn_m = F.n + half_g * (h_m * h_m - hs_m * hs_m);
        f.fn_p = F.n + half_g * (h_p * h_p - hs_p * hs_p);
    }
    return f;
}

struct MCell {
    double w, hu, hv, b;  // w = h, or eta = h + B with bathymetry
};

struct MSlope {
    double w, hu, hv;
};

// Half the limited slope, 0.5 * minmod_scalar(a, b) (reconstruction/muscl.hpp:8-11; the
// states are U +- 0.5 * sigma, :16-36): under the modular range guard every operand is 0
// or at least 2^-152 in magnitude, so the median form of swe_face.h is exact (same bits).
template <bool SPLIT>
__device__ __forceinline__ double half_slope(double a, double b) {
    if constexpr (SPLIT) return half_minmod_split(a, b);
    return 0.5 * minmod(a, b);
}

// The half slopes of cell c from its neighbours m and p along x.
template <bool SPLIT>
__device__ __forceinline__ MSlope half_slope_of(const MCell &m, const MCell &c, const MCell &p) {
    return MSlope{half_slope<SPLIT>(c.w - m.w, p.w - c.w),
                  half_slope<SPLIT>(c.hu - m.hu, p.hu - c.hu),
                  half_slope<SPLIT>(c.hv - m.hv, p.hv - c.hv)};
}

template <int STAGE, int RECON, int RIEMANN, bool BATHY>
__global__ __launch_bounds__(256) void modular_stage_kernel(const GridDev G, const StageArgs A) {
    // scalar loads ahead of the launch's first store: dt and the other stages' run status
    // words (kStatusWords)
    const double dt = STAGE != 0 ? *A.dt : 0.0;
    const int lane = threadIdx.x & (kWave - 1);
    if constexpr (STAGE != 0) {
        if (stage_stop_words<STAGE>(A.flag)) {  // an earlier stage of this run went non-finite
            stage_skip<STAGE>(A.flag);
            return;
        }
        if (STAGE == 1 && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
            atomicAdd(A.flag, 1);
    }
    const int strip = blockIdx.x * kBlockWaves + threadIdx.x / kWave;
    if (strip >= A.nstrips) return;
    const int ng = G.ng, nx = G.nx, ny = G.ny, s = G.stride;
    const int r0 = ng + (int)blockIdx.y * A.tile_rows;
    const int r1 = (r0 + A.tile_rows < ng + nx) ? r0 + A.tile_rows : ng + nx;
    const int rmax = ng + nx + 1;
    const int j = ng - 2 + strip * kStripCols + lane;
    bool gcol;
    const int jc = mirror<0>(j < ng + ny + 2 ? j : ng + ny + 1, ng, ng + ny, gcol);
    unsigned jb = (unsigned)j * 8u, jcb = (unsigned)jc * 8u;
    const bool owner = lane >= 2 && lane < kWave - 2 && j < ng + ny;
    // the y face j+1/2 is one the reference computes (j in [ng-1, ng+ny-1],
    // flux_assembly.hpp:33-34); x faces count in the interior columns only (:16-17)
    const bool yface = lane >= 1 && lane < kWave - 2 && j >= ng - 1 && j <= ng + ny - 1;
    const double g = G.g, half_g = G.half_g;
    int bad = 0;   // HLL without dissipation speed, or a non-finite stage output
    int fbad = 0;  // per face, then masked
    // Wave-uniform, sticky: every cell finished so far by this wave passes the modular range
    // guard, so its faces may use the split division and sqrt forms (all three Riemann
    // solvers: with the guard evaluated off the load path, MUSCL + Rusanov measured 7.5%
    // faster with the split forms and piecewise-constant Rusanov the same).
    bool split = G.split_ok != 0;

    // Row r as stored, with h in the w slot.  Only the loads are issued here: nothing reads
    // the values before `finish`, so the wave keeps computing while they are in flight.
    auto load_raw = [&](int r, bool &gr) {
        pin_lane_offset(jcb, r);
        const int rs = mirror<0>(r, ng, ng + nx, gr);
        const long long ro = __builtin_amdgcn_readfirstlane(rs * s);
        MCell c;
        c.w = ld(row_ptr(A.in_h, ro), jcb);
        c.hu = ld(row_ptr(A.in_hu, ro), jcb);
        c.hv = ld(row_ptr(A.in_hv, ro), jcb);
        c.b = BATHY ? ld(row_ptr(A.B, ro), jcb) : 0.0;
        return c;
    };
    // A loaded row made usable: the reflecting sign flips, w = h + b, and the range guard.
    auto finish = [&](MCell &c, bool gr) {
        const double h = c.w;
        if (gr) c.hu = -c.hu;
        if (gcol) c.hv = -c.hv;
        c.w = BATHY ? h + c.b : h;
        unsigned long long ok = mod_in_range(h) & mod_zero_or_in_range(c.hu) &
                                mod_zero_or_in_range(c.hv);
        if (BATHY) ok &= mod_zero_or_in_range(c.b);
        const unsigned long long exec = __builtin_amdgcn_read_exec();
        split = split && (ok & exec) == exec;
    };
    // x face between rows a (minus, slope sa) and b (plus, slope sb): normal hu
    auto xface = [&](auto split_tag, const MCell &a, const MSlope &sa, const MCell &b,
                     const MSlope &sb) {
        constexpr bool SPLIT = decltype(split_tag)::value;
        NT m{a.w, a.hu, a.hv}, p{b.w, b.hu, b.hv};
        if (RECON == kReconMUSCL) {
            m = NT{a.w + sa.w, a.hu + sa.hu, a.hv + sa.hv};  // sa, sb: half slopes
            p = NT{b.w - sb.w, b.hu - sb.hu, b.hv - sb.hv};
        }
        fbad = 0;
        const FaceFlux f = face_mod<0, RIEMANN, BATHY, SPLIT>(g, half_g, m, a.b, p, b.b, fbad, 0ull);
        bad |= owner ? fbad : 0;
        return f;
    };
    const MSlope zero{0.0, 0.0, 0.0};

    // prologue: rows r0-2 .. r0+2 (all five in flight before the first is used), the plus
    // side of face r0-1/2
    bool gm2, gm1, g0, g1, g2;
    MCell Pm2 = load_raw(r0 - 2, gm2), Pm1 = load_raw(r0 - 1, gm1);
    MCell P0 = load_raw(r0, g0), P1 = load_raw(r0 + 1, g1);
    MCell P2 = load_raw(r0 + 2 <= rmax ? r0 + 2 : rmax, g2);
    finish(Pm2, gm2);
    finish(Pm1, gm1);
    finish(P0, g0);
    finish(P1, g1);
    finish(P2, g2);
    MSlope S0 = zero, S1 = zero;
    if (RECON == kReconMUSCL) S0 = half_slope_of<false>(Pm1, P0, P1);
    const MSlope Sm1 = RECON == kReconMUSCL ? half_slope_of<false>(Pm2, Pm1, P0) : zero;
    FaceFlux Fm;
    Fm = split ? xface(std::true_type{}, Pm1, Sm1, P0, S0)
               : xface(std::false_type{}, Pm1, Sm1, P0, S0);

    // The row's faces (the split and IEEE variants): the slope of row r+1, x face r+1/2 and
    // the y faces of row r.
    auto faces = [&](auto split_tag, const MCell &P0, const MCell &P1, const MCell &P2,
                     const MSlope &S0, MSlope &S1, FaceFlux &Fx, FaceFlux &Fy) {
        constexpr bool SPLIT = decltype(split_tag)::value;
        if (RECON == kReconMUSCL) S1 = half_slope_of<SPLIT>(P0, P1, P2);
        Fx = xface(split_tag, P0, S0, P1, S1);

        // y faces: lane j takes face j+1/2 (normal hv, tangential hu).  The lane moves fill
        // the lane without a source with 0: lane 63's face, lane 62's (it reads lane 63's
        // slope) and lane 0's slope and divergence belong to halo columns, whose results
        // are never stored and whose flags are masked by `yface` / `owner`.
        NT m{P0.w, P0.hv, P0.hu};
        double pw = from_next_z(P0.w), phv = from_next_z(P0.hv), phu = from_next_z(P0.hu);
        if (RECON == kReconMUSCL) {
            const double dw = pw - P0.w, dv = phv - P0.hv, du = phu - P0.hu;
            const double sw = half_slope<SPLIT>(from_prev_z(dw), dw);
            const double sv = half_slope<SPLIT>(from_prev_z(dv), dv);
            const double su = half_slope<SPLIT>(from_prev_z(du), du);
            m = NT{P0.w + sw, P0.hv + sv, P0.hu + su};
            pw = from_next_z(P0.w - sw);
            phv = from_next_z(P0.hv - sv);
            phu = from_next_z(P0.hu - su);
        }
        const double pb = BATHY ? from_next_z(P0.b) : 0.0;
        fbad = 0;
        Fy = face_mod<1, RIEMANN, BATHY, SPLIT>(g, half_g, m, P0.b, NT{pw, phv, phu}, pb, fbad,
                                                kLastLanes);
        bad |= yface ? fbad : 0;
    };
    // The rest of the row, one copy for both face variants (two copies would be merged by
    // the compiler at their common stores, which then take a per-lane 64-bit address):
    // divergence, stage combination, floor, non-finite flag, stores.
    auto tail = [&](int r, const MCell &P0, const FaceFlux &Fm, const FaceFlux &Fx,
                    const FaceFlux &Fy, const NT &u0, double in_h) {
        const double fyp_h = from_prev_z(Fy.fh), fyp_hu = from_prev_z(Fy.ft);
        const double fyp_hv = from_prev_z(Fy.fn_p);

        // finite_volume(_bathy).hpp
        const double rh = divergence(G.inv_dx, G.inv_dy, Fx.fh, Fm.fh, Fy.fh, fyp_h);
        const double rhu = divergence(G.inv_dx, G.inv_dy, Fx.fn_m, Fm.fn_p, Fy.ft, fyp_hu);
        const double rhv = divergence(G.inv_dx, G.inv_dy, Fx.ft, Fm.ft, Fy.fn_m, fyp_hv);

        // The early loads are read here on every path: the stage combination is used only
        // in the owner branch, which the compiler may skip when no lane owns a column, and
        // loads still in flight at the next row would make it wait there for them -- and,
        // the vector memory counter being in order, for this row's stores as well.
        if constexpr (STAGE >= 2) asm volatile("" ::"v"(u0.h), "v"(u0.n), "v"(u0.t));
        if constexpr (STAGE >= 1 && BATHY) asm volatile("" ::"v"(in_h));
        if (owner) {
            const long long ro = __builtin_amdgcn_readfirstlane(r * s);
            pin_lane_offset(jb, r);
            if constexpr (STAGE == 0) {
                st(row_ptr(A.out_h, ro), jb, rh);
                st(row_ptr(A.out_hu, ro), jb, rhu);
                st(row_ptr(A.out_hv, ro), jb, rhv);
            } else {
                // The stage input U_s at an owner cell (interior, never sign-flipped): P0's hu
                // and hv, and h = P0.w without bathymetry; with bathymetry P0.w is eta and h
                // was loaded at the start of the row.
                const double sh = BATHY ? in_h : P0.w;
                double hN, huN, hvN;
                if constexpr (STAGE == 1) {
                    // ssp_rk3.hpp:10-15 (1.0 * U + dt * L)
                    hN = 1.0 * sh + dt * rh;
                    huN = 1.0 * P0.hu + dt * rhu;
                    hvN = 1.0 * P0.hv + dt * rhv;
                } else {
                    const double a = STAGE == 2 ? 3.0 / 4.0 : 1.0 / 3.0;
                    const double b = STAGE == 2 ? 1.0 / 4.0 : 2.0 / 3.0;
                    const double c = STAGE == 2 ? dt / 4.0 : 2.0 * dt / 3.0;
                    hN = a * u0.h + b * sh + c * rh;
                    huN = a * u0.n + b * P0.hu + c * rhu;
                    hvN = a * u0.t + b * P0.hv + c * rhv;
                }
                if constexpr (BATHY) {
                    // serial_solver.hpp:453-472
                    if (hN < kFloor) {
                        hN = kFloor;
                        huN = 0.0;
                        hvN = 0.0;
                    }
                    if (!isfinite(hN) || !isfinite(huN) || !isfinite(hvN)) bad = 1;
                }
                pin_lane_offset(jb, r);
                st(row_ptr(A.out_h, ro), jb, hN);
                st(row_ptr(A.out_hu, ro), jb, huN);
                st(row_ptr(A.out_hv, ro), jb, hvN);
            }
        }
    };
    // One output row: rows r (P0), r+1 (P1), r+2 (P2) are loaded, row r+3 goes into Pn; S0 is
    // the slope of row r, S1 receives that of row r+1; Fm is face r-1/2, Fx receives r+1/2.
    auto row = [&](int r, const MCell &P0, const MCell &P1, const MCell &P2, MCell &Pn,
                   const MSlope &S0, MSlope &S1, const FaceFlux &Fm, FaceFlux &Fx) {
        // row r+3 is first used by the next row: it is finished (and guarded) after this
        // row's faces, a row after its loads were issued
        bool gr;
        Pn = load_raw(r + 3 <= rmax ? r + 3 : rmax, gr);
        // U^n of row r (stages 2, 3) and, with bathymetry, h of row r, for the stage
        // combination: issued now, at the source column (the own column for owner lanes,
        // in bounds for every lane), and read at the end of the row
        NT u0{0.0, 0.0, 0.0};
        double in_h = 0.0;
        if constexpr (STAGE >= 2 || (STAGE == 1 && BATHY)) {
            pin_lane_offset(jcb, r);
            const long long ro = __builtin_amdgcn_readfirstlane(r * s);
            if constexpr (STAGE >= 2)
                u0 = NT{ld(row_ptr(A.u0_h, ro), jcb), ld(row_ptr(A.u0_hu, ro), jcb),
                        ld(row_ptr(A.u0_hv, ro), jcb)};
            if constexpr (BATHY) in_h = ld(row_ptr(A.in_h, ro), jcb);
        }
        FaceFlux Fy;
        if (split)
            faces(std::true_type{}, P0, P1, P2, S0, S1, Fx, Fy);
        else
            faces(std::false_type{}, P0, P1, P2, S0, S1, Fx, Fy);
        tail(r, P0, Fm, Fx, Fy, u0, in_h);
        finish(Pn, gr);
    };

    // march, unrolled 4x so the register window rotates by renaming instead of copies
    MCell P3;
    FaceFlux F1;
    FaceFlux &F0 = Fm;
    for (int r = r0; r < r1; r += 4) {
        row(r, P0, P1, P2, P3, S0, S1, F0, F1);
        if (r + 1 >= r1) break;
        row(r + 1, P1, P2, P3, P0, S1, S0, F1, F0);
        if (r + 2 >= r1) break;
        row(r + 2, P2, P3, P0, P1, S0, S1, F0, F1);
        if (r + 3 >= r1) break;
        row(r + 3, P3, P0, P1, P2, S1, S0, F1, F0);
    }
    if (__any(bad)) {
        // the right-hand side alone (STAGE 0) reports as stage 1, whose first half it is
        if (lane == 0) atomicOr(A.flag + (STAGE == 0 ? 1 : STAGE), kStatusBad);
    }
}

template <int STAGE, int RECON, int RIEMANN, bool BATHY>
hipError_t launch_mod(const GridDev &G, const StageArgs &A, hipStream_t stream, dim3 *grid_out) {
    const dim3 grid = stage_grid(G, A.tile_rows, kBlockWaves);
    if (grid_out) *grid_out = grid;
    hipLaunchKernelGGL((modular_stage_kernel<STAGE, RECON, RIEMANN, BATHY>), grid,
                       dim3(kWave * kBlockWaves), 0, stream, G, A);
    return hipGetLastError();
}

template <int RECON, int RIEMANN, bool BATHY>
hipError_t launch_mod_stage(int stage, const GridDev &G, const StageArgs &A, hipStream_t stream,
                            dim3 *grid_out) {
    switch (stage) {
        case 0: return launch_mod<0, RECON, RIEMANN, BATHY>(G, A, stream, grid_out);
        case 1: return launch_mod<1, RECON, RIEMANN, BATHY>(G, A, stream, grid_out);
        case 2: return launch_mod<2, RECON, RIEMANN, BATHY>(G, A, stream, grid_out);
        case 3: return launch_mod<3, RECON, RIEMANN, BATHY>(G, A, stream, grid_out);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool modular_supported(bool bathy, int recon, int riemann) {
    if (recon != kReconMUSCL && recon != kReconPC) return false;
    if (bathy) return riemann == kHLL;
    return riemann == kHLL || riemann == kRusanov || riemann == kROE;
}

hipError_t launch_modular_stage(int stage, bool bathy, int recon, int riemann, const GridDev &G,
                                const StageArgs &A, hipStream_t stream, dim3 *grid_out) {
    if (G.ng != 2 || A.tile_rows < 1 || A.nstrips != strips_for(G.ny) ||
        !modular_supported(bathy, recon, riemann))
        return hipErrorInvalidValue;
    if (bathy) {
        return recon == kReconMUSCL
                   ? launch_mod_stage<kReconMUSCL, kHLL, true>(stage, G, A, stream, grid_out)
                   : launch_mod_stage<kReconPC, kHLL, true>(stage, G, A, stream, grid_out);
    }
    if (recon == kReconMUSCL) {
        switch (riemann) {
            case kHLL: return launch_mod_stage<kReconMUSCL, kHLL, false>(stage, G, A, stream, grid_out);
            case kRusanov:
                return launch_mod_stage<kReconMUSCL, kRusanov, false>(stage, G, A, stream, grid_out);
            default: return launch_mod_stage<kReconMUSCL, kROE, false>(stage, G, A, stream, grid_out);
        }
    }
    switch (riemann) {
        case kHLL: return launch_mod_stage<kReconPC, kHLL, false>(stage, G, A, stream, grid_out);
        case kRusanov: return launch_mod_stage<kReconPC, kRusanov, false>(stage, G, A, stream, grid_out);
        default: return launch_mod_stage<kReconPC, kROE, false>(stage, G, A, stream, grid_out);
    }
}

}  // namespace swe
