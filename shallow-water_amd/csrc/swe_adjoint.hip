// swe_adjoint.hip -- discrete adjoint of one fused SSP-RK3 stage (gfx950, FP64 VALU).
//
// The forward stage s maps its input V (U^n, U1 or U2) to
//     W = a_s * U^n + c_s * (V + dt * R(V)),   out = floor(W)
// (a, c) = (0, 1), (3/4, 1/4), (1/3, 2/3), with R the MUSCL / hydrostatic-reconstruction /
// HLL right-hand side of include/backend/openmp/omp_kernels.hpp:126-286 and floor the
// positivity fix of :309-331.  Given lam = dL/d(out) this kernel computes, at every interior
// cell,
//     lamW = lam where the output is not the floor state (1e-8, 0, 0), else 0
//     dL/dV      = c_s * lamW + R'(V)^T (c_s * dt * lamW)
//     dL/dU^n   += a_s * lamW                   (a separate accumulator, stages 2 and 3)
//     dL/dB     += R_B'(V)^T (c_s * dt * lamW)  (accumulated over stages and steps)
// with dt frozen (a given number, no derivative).  Stage 1 adds the U^n accumulator to its
// output, so after the three stages of a step the output is dL/dU^n.
//
// Derivative conventions at the non-smooth points (one-sided: the branch the forward took):
//   minmod(a, b)         0 if a*b <= 0, else the argument std::min(|a|, |b|) returned
//                        (|a| on a tie);
//   max(0, x)            1 iff x > 0;
//   std::min / std::max  the argument returned (the first on a tie), for the HLL wave speeds
//                        and b* = max(b-, b+);
//   safe_velocity(q, h)  0 unless h > eps;
//   sqrt(g h)            0 at h = 0;
//   HLL                  the branch taken (left, right, star);
//   floor                0 where the stage output is the floor state.
// The same conventions are used by the CPU dual-number tangent the tests compare against.
//
// Work decomposition (the forward's march, run backwards through each face).  A wave owns
// 58 output columns with a 3-column halo each side and marches down x rows.  Per row it
// recomputes the primal slopes and faces exactly as the forward kernel does (same
// expressions, so every branch decision is the forward's), then reverses each face.  A
// face's reverse yields the adjoints of its two reconstructed states and of the two
// bottoms; the reconstructed state of cell q is x_q +/- sigma_q, so the adjoint splits into
// a direct part for cell q and a slope part sigma_q, which the minmod selection routes to
// q-1, q or q+1.  Along x all of this stays in registers (the accumulator of row q is
// complete once face q+1 is reversed); along y it moves between lanes with the same DPP
// shifts as the forward.  Ghost cells are virtual (mirrored loads, as in the forward), so
// their adjoints are folded back into the cells they mirror, with the reflecting walls'
// sign flips (hu on x ghosts, hv on y ghosts).
//
// Launches: adjoint_stage_kernel does a whole stage in one launch, the x faces of a block
// of rows passing through LDS (the default); adjoint_xface_kernel + adjoint_y_kernel do it
// in two, the x faces passing through device memory.  Both sum every cell's terms in the
// same order, so they give the same bits.
//
// HBM traffic per cell and stage: read V (24 B), B (8), lam (24), the output's floor mask
// (1), write dL/dV (24), the floor mask of V (1), U^n accumulator (24 written; stage 2 also
// reads it, stage 1 only reads it), dL/dB read + write (16): 122 / 146 / 122 bytes for
// stages 3 / 2 / 1.
#include "swe_face.h"
#include "swe_internal.h"
#include "swe_lanes.h"

namespace swe {
namespace {

constexpr int kAdjCols = kWave - 6;  // 58 output columns, 3-column halo on each side

// A cell of V as the faces see it (ghosts mirrored, signs applied); h enters only via eta.
struct PCell {
    double eta, hu, hv, b;
};

// (h, hu, hv) components; for a reconstructed-state slope the first is eta's.
struct D3 {
    double e, u, v;
};

// Adjoint accumulator of one cell: eta (= h's adjoint; B's too), hu, hv, and the part of
// B's adjoint that enters through the bottoms directly.
struct A4 {
    double e, u, v, b;
};

__device__ __forceinline__ bool floor_state(double h, double hu, double hv) {
    return h == kFloor && hu == 0.0 && hv == 0.0;
}

// 0.5 * minmod(a, b) as the forward computes it, and which argument std::min selected:
// 0 none (a*b <= 0), 1 a, 2 b (|b| < |a|).
__device__ __forceinline__ double half_minmod(double a, double b, int &sel) {
    sel = (a * b <= 0.0) ? 0 : (fabs(b) < fabs(a) ? 2 : 1);
    return 0.5 * minmod(a, b);
}

// sigma_q = 0.5*minmod(x_q - x_{q-1}, x_{q+1} - x_q) reversed: the adjoint `sb` of sigma_q
// goes to (q-1, q, q+1) as (-sb/2, sb/2, 0) when a was selected, (0, -sb/2, sb/2) when b.
__device__ __forceinline__ void slope_parts(int sel, double sb, double &to_prev, double &to_self,
                                            double &to_next) {
    const double hsb = 0.5 * sb;
    const double da = sel == 1 ? hsb : 0.0;
    const double db = sel == 2 ? hsb : 0.0;
    to_prev = -da;
    to_self = da - db;
    to_next = db;
}

// Velocities of a reconstructed state exactly as the forward computes them.
// In split mode r receives the refined reciprocal of h the division used (the reverse
// reuses it); otherwise r is left for the reverse to compute.
template <bool SPLIT>
__device__ __forceinline__ void velocities(double n, double t, double h, double &un,
                                           double &ut, double &r) {
    if constexpr (SPLIT) {
        r = div_denominator(h);
        const double qn = div_numerator(n, h, r), qt = div_numerator(t, h, r);
        un = h > kEps ? qn : 0.0;
        ut = h > kEps ? qt : 0.0;
    } else {
        un = safe_velocity(n, h);
        ut = safe_velocity(t, h);
    }
}

// 1/h for the reverse: the primal's reciprocal in split mode, a refined one otherwise.
template <bool SPLIT>
__device__ __forceinline__ double inv_of(double h, double r_primal) {
    if constexpr (SPLIT) return r_primal;
    return recip(h);
}

struct FaceAdj {
    double em, nm, tm, bm;  // adjoints of the minus state (eta, n, t) and its bottom
    double ep, np, tp, bp;  // same for the plus state
};

// Reverse of one face (face_flux + hll of swe_face.h, exact mode).  Inputs: the
// reconstructed minus/plus states as the forward forms them, and the adjoints of the face's
// outputs as seen from the minus cell (mh, mn, mt for fh, fn_m, ft) and the plus cell
// (ph, pn, pt for fh, fn_p, ft).
template <bool SPLIT>
__device__ __forceinline__ FaceAdj face_adjoint(double half_g, double g, double eta_m,
                                                double n_m, double t_m, double b_m,
                                                double eta_p, double n_p, double t_p,
                                                double b_p, double mh, double mn, double mt,
                                                double ph, double pn, double pt) {
    // ---- primal (omp_kernels.hpp:154-189 / :50-86), bitwise the forward's values
    const double b_star = fmax(b_m, b_p);
    const bool bstar_p = b_m < b_p;  // std::max returned its second argument
    const double xm = eta_m - b_m, xp = eta_p - b_p;
    const double h_m = positive_height(xm), h_p = positive_height(xp);
    double un_m, ut_m, un_p, ut_p, r_m = 0.0, r_p = 0.0, r_L = 0.0, r_R = 0.0;
    velocities<SPLIT>(n_m, t_m, h_m, un_m, ut_m, r_m);
    velocities<SPLIT>(n_p, t_p, h_p, un_p, ut_p, r_p);
    const double ym = eta_m - b_star, yp = eta_p - b_star;
    const double hL = positive_height(ym), hR = positive_height(yp);
    const double nL = hL * un_m, tL = hL * ut_m, nR = hR * un_p, tR = hR * ut_p;
    double unL, utL, unR, utR;
    velocities<SPLIT>(nL, tL, hL, unL, utL, r_L);
    velocities<SPLIT>(nR, tR, hR, unR, utR, r_R);
    const double cL = ieee_sqrt<SPLIT>(g * hL), cR = ieee_sqrt<SPLIT>(g * hR);
    const double a1 = unL - cL, a2 = unR - cR, b1 = unL + cL, b2 = unR + cR;
    const double sL = fmin(a1, a2), sR = fmax(b1, b2);
    const bool sL_R = a2 < a1;  // std::min returned its second argument
    const bool sR_R = b1 < b2;  // std::max returned its second argument
    const double FL_h = nL, FL_n = nL * unL + half_g * hL * hL, FL_t = nL * utL;
    const double FR_h = nR, FR_n = nR * unR + half_g * hR * hR, FR_t = nR * utR;

    // ---- reverse
    const double Fh = mh + ph, Fn = mn + pn, Ft = mt + pt;
    // corr = half_g*(h^2 - hs^2) on fn_m / fn_p
    double bh_m = mn * g * h_m, bh_p = pn * g * h_p;
    double bhL = -mn * g * hL, bhR = -pn * g * hR;
    double bnL = 0.0, btL = 0.0, bnR = 0.0, btR = 0.0;
    double gLh = 0.0, gLn = 0.0, gLt = 0.0, gRh = 0.0, gRn = 0.0, gRt = 0.0;
    double bsL = 0.0, bsR = 0.0;
    if (sL >= 0.0) {
        gLh = Fh;
        gLn = Fn;
        gLt = Ft;
    } else if (sR <= 0.0) {
        gRh = Fh;
        gRn = Fn;
        gRt = Ft;
    } else {
        // F = (sR*FL - sL*FR + sL*sR*(UR - UL)) / (sR - sL)
        const double inv = recip(sR - sL);
        const double dh = hR - hL, dn = nR - nL, dtt = tR - tL, sLR = sL * sR;
        const double Sh = (sR * FL_h - sL * FR_h + sLR * dh) * inv;
        const double Sn = (sR * FL_n - sL * FR_n + sLR * dn) * inv;
        const double St = (sR * FL_t - sL * FR_t + sLR * dtt) * inv;
        const double Nh = Fh * inv, Nn = Fn * inv, Nt = Ft * inv;
        const double bD = -(Fh * Sh + Fn * Sn + Ft * St) * inv;
        bsR = bD + Nh * (FL_h + sL * dh) + Nn * (FL_n + sL * dn) + Nt * (FL_t + sL * dtt);
        bsL = -bD + Nh * (sR * dh - FR_h) + Nn * (sR * dn - FR_n) + Nt * (sR * dtt - FR_t);
        gLh = Nh * sR;
        gLn = Nn * sR;
        gLt = Nt * sR;
        gRh = -Nh * sL;
        gRn = -Nn * sL;
        gRt = -Nt * sL;
        bhR += Nh * sLR;
        bhL -= Nh * sLR;
        bnR += Nn * sLR;
        bnL -= Nn * sLR;
        btR += Nt * sLR;
        btL -= Nt * sLR;
    }
    // FL = (nL, nL*unL + half_g*hL^2, nL*utL), FR likewise
    bnL += gLh + gLn * unL + gLt * utL;
    double bunL = gLn * nL, butL = gLt * nL;
    bhL += gLn * g * hL;
    bnR += gRh + gRn * unR + gRt * utR;
    double bunR = gRn * nR, butR = gRt * nR;
    bhR += gRn * g * hR;
    // wave speeds
    double bcL = 0.0, bcR = 0.0;
    if (sR_R) {
        bunR += bsR;
        bcR += bsR;
    } else {
        bunL += bsR;
        bcL += bsR;
    }
    if (sL_R) {
        bunR += bsL;
        bcR -= bsL;
    } else {
        bunL += bsL;
        bcL -= bsL;
    }
    // c = sqrt(g*h): dc/dh = g / (2c), 0 at h = 0
    if (cL > 0.0) bhL += bcL * (0.5 * g) * recip(cL);
    if (cR > 0.0) bhR += bcR * (0.5 * g) * recip(cR);
    // HLL velocities
    if (hL > kEps) {
        const double r = inv_of<SPLIT>(hL, r_L);
        bnL += bunL * r;
        btL += butL * r;
        bhL -= (bunL * unL + butL * utL) * r;
    }
    if (hR > kEps) {
        const double r = inv_of<SPLIT>(hR, r_R);
        bnR += bunR * r;
        btR += butR * r;
        bhR -= (bunR * unR + butR * utR) * r;
    }
    // HLL inputs: (hs, hs*u, hs*v)
    bhL += bnL * un_m + btL * ut_m;
    const double bun_m = bnL * hL, but_m = btL * hL;
    bhR += bnR * un_p + btR * ut_p;
    const double bun_p = bnR * hR, but_p = btR * hR;
    // hs = max(0, eta - b*)
    FaceAdj r{};
    double bbs = 0.0;
    if (ym > 0.0) {
        r.em += bhL;
        bbs -= bhL;
    }
    if (yp > 0.0) {
        r.ep += bhR;
        bbs -= bhR;
    }
    // face velocities u = n / h
    if (h_m > kEps) {
        const double q = inv_of<SPLIT>(h_m, r_m);
        r.nm = bun_m * q;
        r.tm = but_m * q;
        bh_m -= (bun_m * un_m + but_m * ut_m) * q;
    }
    if (h_p > kEps) {
        const double q = inv_of<SPLIT>(h_p, r_p);
        r.np = bun_p * q;
        r.tp = but_p * q;
        bh_p -= (bun_p * un_p + but_p * ut_p) * q;
    }
    // h = max(0, eta - b)
    if (xm > 0.0) {
        r.em += bh_m;
        r.bm -= bh_m;
    }
    if (xp > 0.0) {
        r.ep += bh_p;
        r.bp -= bh_p;
    }
    if (bstar_p)
        r.bp += bbs;
    else
        r.bm += bbs;
    return r;
}

__device__ __forceinline__ FaceAdj face_zero() { return FaceAdj{0, 0, 0, 0, 0, 0, 0, 0}; }

// Reflecting walls flip the normal momentum of a ghost: hu across x walls (rows), hv across
// y walls (columns).
template <int BC>
__device__ __forceinline__ void fold_row(A4 &dst, const A4 &ghost) {
    dst.e += ghost.e;
    dst.u += BC == 0 ? -ghost.u : ghost.u;
    dst.v += ghost.v;
    dst.b += ghost.b;
}

// A wave over a strip of 58 output columns (lanes 3..60) with a 3-column halo on each
// side: lane l holds column j0 - 3 + l.  Loads mirror ghost rows and columns (with the
// reflecting walls' sign flips) exactly as the forward stage kernel does.
template <int BC>
struct Lanes {
    const AdjArgs &A;
    int s, ng, nx, ny, lane, j0;
    // own column / mirrored source column, in bytes; re-pinned per row (pin_lane_offset)
    mutable unsigned j_bytes, jc_bytes;
    bool gcol;   // the lane's column is outside the interior (its values are mirrored)
    bool owner;  // the lane writes its column
    bool xact;   // interior column: x faces live, lamW nonzero
    bool yact;   // the y face on the lane's upper side (j | j+1) is in the domain

    __device__ __forceinline__ Lanes(const GridDev &G, const AdjArgs &A_, int strip, int lane_)
        : A(A_), s(G.stride), ng(G.ng), nx(G.nx), ny(G.ny), lane(lane_) {
        j0 = ng + strip * kAdjCols;
        const int j = j0 - 3 + lane;
        const int jcl = j < ng - 2 ? ng - 2 : (j > ng + ny + 1 ? ng + ny + 1 : j);
        const int jc = mirror<BC>(jcl, ng, ng + ny, gcol);
        j_bytes = (unsigned)(j < 0 ? 0 : j) * 8u;
        jc_bytes = (unsigned)jc * 8u;
        gcol = j < ng || j >= ng + ny;
        owner = lane >= 3 && lane < 3 + kAdjCols && j < ng + ny;
        xact = j >= ng && j < ng + ny;
        yact = j >= ng - 1 && j <= ng + ny - 1;
    }

    __device__ __forceinline__ long long row_off(int r) const {
        return (long long)__builtin_amdgcn_readfirstlane(r * s);
    }

    // Row r of V (r within 2 of the interior) as stored, h in the eta slot.  Only the loads
    // are issued here: nothing reads the values before `finish`, so the wave keeps
    // computing while they are in flight.
    __device__ __forceinline__ PCell load_raw(int r, bool &ghost_row) const {
        pin_lane_offset(jc_bytes, r);
        const int rs = mirror<BC>(r, ng, ng + nx, ghost_row);
        const long long ro = row_off(rs);
        PCell c;
        c.eta = ld(row_ptr(A.v_h, ro), jc_bytes);
        c.hu = ld(row_ptr(A.v_hu, ro), jc_bytes);
        c.hv = ld(row_ptr(A.v_hv, ro), jc_bytes);
        c.b = ld(row_ptr(A.B, ro), jc_bytes);
        return c;
    }
    // A loaded row made usable: `floor` = the stored cell is the floor state; the reflecting
    // sign flips; eta = h + b.
    __device__ __forceinline__ void finish(PCell &c, bool ghost_row, bool &floor) const {
        const double h = c.eta;
        floor = floor_state(h, c.hu, c.hv);
        if (BC == 0) {
            if (ghost_row) c.hu = -c.hu;
            if (gcol) c.hv = -c.hv;
        }
        c.eta = h + c.b;
    }
    __device__ __forceinline__ PCell load(int r, bool &floor) const {
        bool gr;
        PCell c = load_raw(r, gr);
        finish(c, gr, floor);
        return c;
    }

    // lamW of row r as stored (r in [ng-2, ng+nx+1], inside the padded layout), with the
    // output's floor flag.  Every lane loads, at the source column (its own column where
    // xact, in bounds everywhere), and rows outside the interior load too: no branch, so
    // nothing forces the values to be read before `finish_lam`.
    struct LamRaw {
        D3 l;
        unsigned m;
        bool row_in;  // wave-uniform: r is an interior row
    };
    __device__ __forceinline__ LamRaw load_lam_raw(int r) const {
        const long long ro = row_off(r);
        pin_lane_offset(jc_bytes, r);
        return LamRaw{D3{ld(row_ptr(A.lam_h, ro), jc_bytes), ld(row_ptr(A.lam_hu, ro), jc_bytes),
                         ld(row_ptr(A.lam_hv, ro), jc_bytes)},
                      A.mask_out[ro + (jc_bytes >> 3)], r >= ng && r < ng + nx};
    }
    // lamW: the output adjoint where the output is not the floor state, 0 outside the
    // interior rows and columns.
    __device__ __forceinline__ D3 finish_lam(const LamRaw &q) const {
        const bool z = !q.row_in || !xact || q.m != 0;
        return D3{z ? 0.0 : q.l.e, z ? 0.0 : q.l.u, z ? 0.0 : q.l.v};
    }
    __device__ __forceinline__ D3 load_lam(int r) const { return finish_lam(load_lam_raw(r)); }

    // The lanes whose cell is inside the split forms' range guard (depth h = eta - b,
    // momenta, bottom), one ballot per compare.
    __device__ __forceinline__ static unsigned long long split_lanes(const PCell &c) {
        return split_ok_lanes(c.eta - c.b) & split_ok_lanes(c.hu) & split_ok_lanes(c.hv) &
               split_ok_lanes(c.b);
    }
    __device__ __forceinline__ static bool all_lanes(unsigned long long m) {
        const unsigned long long exec = __builtin_amdgcn_read_exec();
        return (m & exec) == exec;
    }

    // y faces of one row: the adjoint contributions to the row's cells, ghost columns
    // folded into the columns they mirror.
    template <bool SPLIT>
    __device__ __forceinline__ A4 y_row(double half_g, double g, double ky, const PCell &P,
                                        const D3 &L) const {
        // The lane moves fill the lane without a source (63 resp. 0) with 0.  Those lanes are
        // in the 3-column halo, like every lane their values reach, and no vote reads them.
        const double de = from_next_z(P.eta) - P.eta;
        const double du = from_next_z(P.hu) - P.hu;
        const double dv = from_next_z(P.hv) - P.hv;
        int se_, su_, sv_;
        const double se = half_minmod(from_prev_z(de), de, se_);
        const double su = half_minmod(from_prev_z(du), du, su_);
        const double sv = half_minmod(from_prev_z(dv), dv, sv_);
        const double pe = from_next_z(P.eta - se);
        const double pu = from_next_z(P.hu - su);
        const double pv = from_next_z(P.hv - sv);
        const double pb = from_next_z(P.b);
        const double le = from_next_z(L.e), lu = from_next_z(L.u), lv = from_next_z(L.v);
        // y faces: normal momentum hv, tangential hu
        FaceAdj f = face_adjoint<SPLIT>(half_g, g, P.eta + se, P.hv + sv, P.hu + su, P.b, pe, pv,
                                        pu, pb, -ky * L.e, -ky * L.v, -ky * L.u, ky * le,
                                        ky * lv, ky * lu);
        if (!yact) f = face_zero();
        // plus-side parts of the face below (lane j-1) belong to this lane's cell
        const double qe = from_prev_z(f.ep), qn = from_prev_z(f.np), qt = from_prev_z(f.tp);
        const double qb = from_prev_z(f.bp);
        // slope adjoints of this cell (the minus state uses +sigma, the plus state -sigma),
        // routed by the minmod selection: own part, lane j+1's "previous" part and lane
        // j-1's "next" part land here
        A4 acc;
        double tp, ts, tn;
        slope_parts(se_, f.em - qe, tp, ts, tn);
        acc.e = f.em + qe + ts + from_next_z(tp) + from_prev_z(tn);
        slope_parts(su_, f.tm - qt, tp, ts, tn);
        acc.u = f.tm + qt + ts + from_next_z(tp) + from_prev_z(tn);
        slope_parts(sv_, f.nm - qn, tp, ts, tn);
        acc.v = f.nm + qn + ts + from_next_z(tp) + from_prev_z(tn);
        acc.b = f.bm + qb;
        fold_cols(acc);
        return acc;
    }

    // Ghost columns (y walls) fold into the columns they mirror, hv negated for reflecting.
    __device__ __forceinline__ void fold_cols(A4 &acc) const {
        const int base = j0 - 3;  // lane of column c is c - base
        if (j0 == ng) {           // left wall in this strip: ghosts ng-1, ng-2
            fold_col(acc, ng - 1 - base, ng - base);
            fold_col(acc, ng - 2 - base, (BC == 0 ? ng + 1 : ng) - base);
        }
        const int gr = ng + ny - base;  // lane of ghost column ng+ny
        if (gr <= kWave - 1) {
            fold_col(acc, gr, gr - 1);
            if (gr + 1 <= kWave - 1) fold_col(acc, gr + 1, BC == 0 ? gr - 2 : gr - 1);
        }
    }

    __device__ __forceinline__ void fold_col(A4 &acc, int src, int dst) const {
        const double e = read_lane(acc.e, src), u = read_lane(acc.u, src);
        const double v = read_lane(acc.v, src), b = read_lane(acc.b, src);
        if (lane == dst) {
            acc.e += e;
            acc.u += u;
            acc.v += BC == 0 ? -v : v;
            acc.b += b;
        }
    }
};

template <int STAGE>
__device__ __forceinline__ double stage_c() {
    return STAGE == 1 ? 1.0 : (STAGE == 2 ? 0.25 : 2.0 / 3.0);
}

#ifndef SWE_ADJ_WAVES_PER_EU
#define SWE_ADJ_WAVES_PER_EU 2
#endif

// ---- x gather ------------------------------------------------------------------------
// Sums the x faces' adjoint parts of a row r (rows ng-2 .. ng+nx+1; faces outside the
// domain contribute nothing): the direct parts of faces r (minus side) and r-1 (plus side)
// and the slopes of rows r-1, r, r+1 routed by their minmod selections.  S supplies the
// face outputs xf(k, f) (k = em, nm, tm, bm, ep, np, tp, bp) and the selections sel(t);
// the summation order is fixed here, so every source gives the same bits.
template <class S>
__device__ __forceinline__ A4 gather_row(const S &X, const GridDev &G, int r) {
    auto face_ok = [&](int f) { return f >= G.ng - 1 && f <= G.ng + G.nx - 1; };
    A4 T{0.0, 0.0, 0.0, 0.0};
    if (face_ok(r)) {
        T.e += X.xf(0, r);
        T.u += X.xf(1, r);
        T.v += X.xf(2, r);
        T.b += X.xf(3, r);
    }
    if (face_ok(r - 1)) {
        T.e += X.xf(4, r - 1);
        T.u += X.xf(5, r - 1);
        T.v += X.xf(6, r - 1);
        T.b += X.xf(7, r - 1);
    }
    for (int t = r - 1; t <= r + 1; ++t) {
        if (t < G.ng - 1 || t > G.ng + G.nx) continue;
        // slope adjoint of row t: +minus state of face t, -plus state of face t-1
        double ge = 0.0, gu = 0.0, gv = 0.0;
        if (face_ok(t)) {
            ge += X.xf(0, t);
            gu += X.xf(1, t);
            gv += X.xf(2, t);
        }
        if (face_ok(t - 1)) {
            ge -= X.xf(4, t - 1);
            gu -= X.xf(5, t - 1);
            gv -= X.xf(6, t - 1);
        }
        const int sl = X.sel(t);
        double tp, ts, tn;
        slope_parts(sl & 3, ge, tp, ts, tn);
        T.e += t == r - 1 ? tn : (t == r ? ts : tp);
        slope_parts((sl >> 2) & 3, gu, tp, ts, tn);
        T.u += t == r - 1 ? tn : (t == r ? ts : tp);
        slope_parts((sl >> 4) & 3, gv, tp, ts, tn);
        T.v += t == r - 1 ? tn : (t == r ? ts : tp);
    }
    return T;
}

// x-face outputs in device memory (written by adjoint_xface_kernel).
struct XGather {
    const AdjArgs &A;
    int ng, stride, j;
    long long plane;
    __device__ __forceinline__ double xf(int k, int f) const {
        return A.xf[k * plane + (long long)(f - (ng - 1)) * stride + j];
    }
    __device__ __forceinline__ int sel(int t) const {
        return A.xsel[(long long)(t - (ng - 1)) * stride + j];
    }
};

// x-face outputs of a block's faces F0 .. F1 in LDS (adjoint_stage_kernel): face f at
// [(f - F0) * 8 + k][lane], the selections of rows F0 .. F1+1 after them.
struct LdsGather {
    const double *faces;
    const unsigned char *sels;
    int F0, lane;
    __device__ __forceinline__ double xf(int k, int f) const {
        return faces[((f - F0) * 8 + k) * kWave + lane];
    }
    __device__ __forceinline__ int sel(int t) const { return sels[(t - F0) * kWave + lane]; }
};

// The ghost rows' x parts folded into the interior rows they mirror (reflecting: ng-1 ->
// ng, ng-2 -> ng+1; transmissive: both -> ng; likewise at the far wall).
template <int BC, class S>
__device__ __forceinline__ A4 fold_ghost_rows(const S &X, const GridDev &G, int r, A4 T) {
    const int ng = G.ng, nx = G.nx;
    if (r == ng) {
        fold_row<BC>(T, gather_row(X, G, ng - 1));
        if (BC != 0) fold_row<BC>(T, gather_row(X, G, ng - 2));
    }
    if (BC == 0 && r == ng + 1) fold_row<BC>(T, gather_row(X, G, ng - 2));
    if (r == ng + nx - 1) {
        fold_row<BC>(T, gather_row(X, G, ng + nx));
        if (BC != 0) fold_row<BC>(T, gather_row(X, G, ng + nx + 1));
    }
    if (BC == 0 && r == ng + nx - 2) fold_row<BC>(T, gather_row(X, G, ng + nx + 1));
    return T;
}
template <int BC, class S>
__device__ __forceinline__ A4 gather_folded(const S &X, const GridDev &G, int r) {
    return fold_ghost_rows<BC>(X, G, r, gather_row(X, G, r));
}

// ---- rolling x gather (the fused kernel's rows) ----------------------------------------
// A wave walking rows ra .. rb-1 routes each row's slope adjoint once and keeps the parts
// of rows r-1, r, r+1 in registers, where gather_row routes all three for every row.  The
// sums are gather_row's, term by term and in its order, so the bits are the same.  The LDS
// reads of row r+1's slope inputs are issued before the row's y faces and used after them
// (issuing the direct parts early as well measured no better).  Interior rows r only need slope
// rows ng-1 .. ng+nx, which all exist; a face index is clamped into the block's range
// [F0, F1], which holds every face inside the domain, and a face outside the domain is
// read there but not used.
struct SlopeIn {
    double a[3];  // minus-state adjoints (e, u, v) of face t
    double m[3];  // plus-state adjoints of face t-1
    int sel;
};
struct SlopeRow {
    double p[3], s[3], n[3];  // parts routed to rows t-1, t, t+1 (e, u, v)
};
struct DirectIn {
    double m[4];  // face r, minus side (e, u, v, b)
    double p[4];  // face r-1, plus side
};
__device__ __forceinline__ int clamp_face(int f, int F0, int F1) {
    return f < F0 ? F0 : (f > F1 ? F1 : f);
}
__device__ __forceinline__ SlopeIn slope_in(const LdsGather &X, int F1, int t) {
    const int fa = clamp_face(t, X.F0, F1), fb = clamp_face(t - 1, X.F0, F1);
    SlopeIn I;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        I.a[k] = X.xf(k, fa);
        I.m[k] = X.xf(4 + k, fb);
    }
    I.sel = X.sel(t);
    return I;
}
__device__ __forceinline__ SlopeRow slope_row(const SlopeIn &I, const GridDev &G, int t) {
    double ge = 0.0, gu = 0.0, gv = 0.0;
    if (t >= G.ng - 1 && t <= G.ng + G.nx - 1) {
        ge += I.a[0];
        gu += I.a[1];
        gv += I.a[2];
    }
    if (t - 1 >= G.ng - 1 && t - 1 <= G.ng + G.nx - 1) {
        ge -= I.m[0];
        gu -= I.m[1];
        gv -= I.m[2];
    }
    SlopeRow S;
    slope_parts(I.sel & 3, ge, S.p[0], S.s[0], S.n[0]);
    slope_parts((I.sel >> 2) & 3, gu, S.p[1], S.s[1], S.n[1]);
    slope_parts((I.sel >> 4) & 3, gv, S.p[2], S.s[2], S.n[2]);
    return S;
}
__device__ __forceinline__ DirectIn direct_in(const LdsGather &X, int F1, int r) {
    const int fa = clamp_face(r, X.F0, F1), fb = clamp_face(r - 1, X.F0, F1);
    DirectIn D;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        D.m[k] = X.xf(k, fa);
        D.p[k] = X.xf(4 + k, fb);
    }
    return D;
}
// gather_row(X, G, r) from the direct parts and the slope rows r-1 (Sm), r (S0), r+1 (Sp).
__device__ __forceinline__ A4 gather_rolled(const GridDev &G, int r, const DirectIn &D,
                                           const SlopeRow &Sm, const SlopeRow &S0,
                                           const SlopeRow &Sp) {
    A4 T{0.0, 0.0, 0.0, 0.0};
    if (r >= G.ng - 1 && r <= G.ng + G.nx - 1) {
        T.e += D.m[0];
        T.u += D.m[1];
        T.v += D.m[2];
        T.b += D.m[3];
    }
    if (r - 1 >= G.ng - 1 && r - 1 <= G.ng + G.nx - 1) {
        T.e += D.p[0];
        T.u += D.p[1];
        T.v += D.p[2];
        T.b += D.p[3];
    }
    T.e += Sm.n[0];
    T.u += Sm.n[1];
    T.v += Sm.n[2];
    T.e += S0.s[0];
    T.u += S0.s[1];
    T.v += S0.s[2];
    T.e += Sp.p[0];
    T.u += Sp.p[1];
    T.v += Sp.p[2];
    return T;
}

// The values emit_cell reads back at row r: the U^n accumulator (stages 2, 1) and dL/dB,
// loaded at the source column (the own column for owner lanes, in bounds for every lane)
// so that they can be issued at the start of the row.
struct EmitIn {
    D3 e;
    double lb;
};
template <int STAGE, int BC>
__device__ __forceinline__ EmitIn load_emit(const AdjArgs &A, const Lanes<BC> &w, int r) {
    const long long ro = w.row_off(r);
    unsigned jc = w.jc_bytes;
    pin_lane_offset(jc, r);
    EmitIn E{{0.0, 0.0, 0.0}, 0.0};
    if constexpr (STAGE != 3)
        E.e = D3{ld(row_ptr(A.e_h, ro), jc), ld(row_ptr(A.e_hu, ro), jc),
                 ld(row_ptr(A.e_hv, ro), jc)};
    E.lb = ld(row_ptr(A.lam_b, ro), jc);
    return E;
}
// Reads the early loads on every path through the row: emit_cell runs only in the owner
// branch, which the compiler may skip when no lane owns a column, and loads still in flight
// at the next row would make it wait there -- and, the vector memory counter being in
// order, for this row's stores as well.
template <int STAGE>
__device__ __forceinline__ void use_emit(const EmitIn &E) {
    if constexpr (STAGE != 3) asm volatile("" ::"v"(E.e.e), "v"(E.e.u), "v"(E.e.v));
    asm volatile("" ::"v"(E.lb));
}

// The stage's outputs at an owned cell of row r: dL/dV (stage 1: plus the U^n accumulator),
// the U^n accumulator, dL/dB.  acc = the y faces' part, T = the x faces' part, E = the
// accumulator and dL/dB as they were before this stage.
template <int STAGE, int BC>
__device__ __forceinline__ void emit_cell(const AdjArgs &A, const Lanes<BC> &w, int r,
                                          const D3 &L, const A4 &acc, const A4 &T,
                                          const EmitIn &E) {
    constexpr double cv = STAGE == 1 ? 1.0 : (STAGE == 2 ? 0.25 : 2.0 / 3.0);
    const long long ro = w.row_off(r);
    unsigned jb = w.j_bytes;
    pin_lane_offset(jb, r);
    double oh = cv * L.e + (T.e + acc.e);
    double ou = cv * L.u + (T.u + acc.u);
    double ov = cv * L.v + (T.v + acc.v);
    if constexpr (STAGE == 3) {
        st(row_ptr(A.e_h, ro), jb, (1.0 / 3.0) * L.e);
        st(row_ptr(A.e_hu, ro), jb, (1.0 / 3.0) * L.u);
        st(row_ptr(A.e_hv, ro), jb, (1.0 / 3.0) * L.v);
    } else if constexpr (STAGE == 2) {
        st(row_ptr(A.e_h, ro), jb, E.e.e + 0.75 * L.e);
        st(row_ptr(A.e_hu, ro), jb, E.e.u + 0.75 * L.u);
        st(row_ptr(A.e_hv, ro), jb, E.e.v + 0.75 * L.v);
    } else {
        oh += E.e.e;
        ou += E.e.u;
        ov += E.e.v;
    }
    st(row_ptr(A.out_h, ro), jb, oh);
    st(row_ptr(A.out_hu, ro), jb, ou);
    st(row_ptr(A.out_hv, ro), jb, ov);
    st(row_ptr(A.lam_b, ro), jb, E.lb + ((T.e + T.b) + (acc.e + acc.b)));
}

// ---- y faces, x gather and the stage's outputs: every row independently -----------------
// Runs after adjoint_xface_kernel.  Per row: the y faces' adjoint (lanes, DPP), the x
// faces' parts gathered from xf/xsel (ghost rows folded at the x walls), then the stage's
// outputs: dL/dV (stage 1: plus the U^n accumulator), the U^n accumulator, dL/dB.  A wave
// takes `y_rows` consecutive rows, issuing the next row's loads before each row's
// arithmetic.
template <int STAGE, int BC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SWE_ADJ_WAVES_PER_EU)))
void adjoint_y_kernel(const GridDev G, const AdjArgs A) {
    const int lane = threadIdx.x & (kWave - 1);
    const int strip = blockIdx.x * kBlockWaves + threadIdx.x / kWave;
    if (strip >= A.nstrips) return;
    const Lanes<BC> w(G, A, strip, lane);
    const int ng = G.ng, nx = G.nx;
    const int r0 = ng + (int)blockIdx.y * A.y_rows;
    const int r1 = (r0 + A.y_rows < ng + nx) ? r0 + A.y_rows : ng + nx;
    constexpr double cv = STAGE == 1 ? 1.0 : (STAGE == 2 ? 0.25 : 2.0 / 3.0);
    const double ky = G.inv_dy * (cv * *A.dt);
    const XGather X{A, ng, G.stride, (int)(w.j_bytes >> 3), (long long)(nx + 1) * G.stride};
    bool fl;
    PCell P = w.load(r0, fl);
    D3 L = w.load_lam(r0);
#pragma unroll 1
    for (int r = r0; r < r1; ++r) {
        // the next row, its lamW and this row's read-back values are loaded first and
        // finished after the row's arithmetic
        const int rn = r + 1 < r1 ? r + 1 : r;
        bool gr;
        PCell Pn = w.load_raw(rn, gr);
        const auto Lnr = w.load_lam_raw(rn);
        const EmitIn E = load_emit<STAGE, BC>(A, w, r);
        A4 acc = (G.split_ok && Lanes<BC>::all_lanes(Lanes<BC>::split_lanes(P)))
                     ? w.template y_row<true>(G.half_g, G.g, ky, P, L)
                     : w.template y_row<false>(G.half_g, G.g, ky, P, L);
        use_emit<STAGE>(E);
        if (w.owner) emit_cell<STAGE, BC>(A, w, r, L, acc, gather_folded<BC>(X, G, r), E);
        w.finish(Pn, gr, fl);
        P = Pn;
        L = w.finish_lam(Lnr);
    }
}

// ---- x faces: reverse each face f | f+1 independently -------------------------------------
// Faces f = ng-1 .. ng+nx-1 (index i = f - (ng-1)).  A wave takes `tile_rows` consecutive
// faces of its strip, carrying only the primal rows (each row loaded once per wave) and
// the slopes.  Per face it writes the 8 adjoint outputs (eta, hu, hv, bottom of the minus
// and plus states) to xf[k][i][j], and per row q = f the minmod selections of its slope to
// sel[q - (ng-1)][j] (plus row ng+nx from the last face).  It also writes the floor mask
// of V for the rows it owns (q = f, interior).
template <int STAGE, int BC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SWE_ADJ_WAVES_PER_EU)))
void adjoint_xface_kernel(const GridDev G, const AdjArgs A) {
    const int lane = threadIdx.x & (kWave - 1);
    const int strip = blockIdx.x * kBlockWaves + threadIdx.x / kWave;
    if (strip >= A.nstrips) return;
    const Lanes<BC> w(G, A, strip, lane);
    const int ng = G.ng, nx = G.nx;
    const int f0 = ng - 1 + (int)blockIdx.y * A.tile_rows;
    const int f1 = (f0 + A.tile_rows < ng + nx) ? f0 + A.tile_rows : ng + nx;
    const double kx = G.inv_dx * (stage_c<STAGE>() * *A.dt);
    const long long plane = (long long)(nx + 1) * G.stride;  // one component of xf
    bool split = G.split_ok != 0;
    bool fl;
    // floor mask of V for the rows this wave owns (q = f for its faces, interior q)
    auto mark = [&](int r, bool floor) {
        if (r >= f0 && r < f1 && r >= ng && r < ng + nx && w.owner)
            A.mask_in[w.row_off(r) + (w.j_bytes >> 3)] = floor ? 1 : 0;
    };

    PCell P0, P1, P2, P3;
    D3 S0, S1;
    int sel0, sel1;
    {
        const PCell Pm = w.load(f0 - 1, fl);
        split = split && Lanes<BC>::all_lanes(Lanes<BC>::split_lanes(Pm));
        P0 = w.load(f0, fl);
        mark(f0, fl);
        P1 = w.load(f0 + 1, fl);
        mark(f0 + 1, fl);
        P2 = w.load(f0 + 2, fl);
        mark(f0 + 2, fl);
        split = split && Lanes<BC>::all_lanes(Lanes<BC>::split_lanes(P0) &
                                                   Lanes<BC>::split_lanes(P1) &
                                                   Lanes<BC>::split_lanes(P2));
        int sa, sb, sc;
        S0.e = half_minmod(P0.eta - Pm.eta, P1.eta - P0.eta, sa);
        S0.u = half_minmod(P0.hu - Pm.hu, P1.hu - P0.hu, sb);
        S0.v = half_minmod(P0.hv - Pm.hv, P1.hv - P0.hv, sc);
        sel0 = sa | (sb << 2) | (sc << 4);
    }
    D3 L0 = w.load_lam(f0), L1 = w.load_lam(f0 + 1), L2;

#pragma unroll 1
    for (int f = f0; f < f1; ++f) {
        // next face's row of V (written to the floor mask when owned) and lamW
        // (loaded now, finished after the face)
        const int rn = f + 3 <= ng + nx + 1 ? f + 3 : ng + nx + 1;
        bool gr3;
        P3 = w.load_raw(rn, gr3);
        const auto L2r = w.load_lam_raw(f + 2);

        int sa, sb, sc;
        S1.e = half_minmod(P1.eta - P0.eta, P2.eta - P1.eta, sa);
        S1.u = half_minmod(P1.hu - P0.hu, P2.hu - P1.hu, sb);
        S1.v = half_minmod(P1.hv - P0.hv, P2.hv - P1.hv, sc);
        sel1 = sa | (sb << 2) | (sc << 4);

        // x faces: normal momentum hu, tangential hv
        FaceAdj x;
        if (split)
            x = face_adjoint<true>(G.half_g, G.g, P0.eta + S0.e, P0.hu + S0.u, P0.hv + S0.v,
                                   P0.b, P1.eta - S1.e, P1.hu - S1.u, P1.hv - S1.v, P1.b,
                                   -kx * L0.e, -kx * L0.u, -kx * L0.v, kx * L1.e, kx * L1.u,
                                   kx * L1.v);
        else
            x = face_adjoint<false>(G.half_g, G.g, P0.eta + S0.e, P0.hu + S0.u, P0.hv + S0.v,
                                    P0.b, P1.eta - S1.e, P1.hu - S1.u, P1.hv - S1.v, P1.b,
                                    -kx * L0.e, -kx * L0.u, -kx * L0.v, kx * L1.e, kx * L1.u,
                                    kx * L1.v);
        if (w.owner) {
            const long long fo = (long long)__builtin_amdgcn_readfirstlane((f - (ng - 1)) * w.s);
            double *xf = A.xf + fo;
            st(row_ptr(xf, 0 * plane), w.j_bytes, x.em);
            st(row_ptr(xf, 1 * plane), w.j_bytes, x.nm);
            st(row_ptr(xf, 2 * plane), w.j_bytes, x.tm);
            st(row_ptr(xf, 3 * plane), w.j_bytes, x.bm);
            st(row_ptr(xf, 4 * plane), w.j_bytes, x.ep);
            st(row_ptr(xf, 5 * plane), w.j_bytes, x.np);
            st(row_ptr(xf, 6 * plane), w.j_bytes, x.tp);
            st(row_ptr(xf, 7 * plane), w.j_bytes, x.bp);
            unsigned char *sr = A.xsel + fo;
            sr[w.j_bytes >> 3] = (unsigned char)sel0;
            if (f == ng + nx - 1) sr[w.s + (w.j_bytes >> 3)] = (unsigned char)sel1;
        }
        w.finish(P3, gr3, fl);
        if (rn == f + 3) mark(rn, fl);
        L2 = w.finish_lam(L2r);
        split = split && Lanes<BC>::all_lanes(Lanes<BC>::split_lanes(P3));
        P0 = P1;
        P1 = P2;
        P2 = P3;
        S0 = S1;
        sel0 = sel1;
        L0 = L1;
        L1 = L2;
    }
}

// ---- the whole adjoint stage in one launch --------------------------------------------
// A block of W waves covers one strip and the rows R0 .. R0 + W*T - 1 (T = tile_rows).
// Phase 1: the block's x faces F0 = max(R0-2, ng-1) .. F1 = min(R0 + W*T, ng+nx-1) -- the
// faces its rows' x parts depend on -- are split into W contiguous runs, one per wave; each
// wave marches its run as adjoint_xface_kernel does and writes every face's 8 outputs and
// each row's minmod selections to LDS.  Phase 2, after the barrier: each wave takes T rows,
// reverses their y faces (lanes, DPP), gathers the x parts from LDS with the ghost-row
// folds, and writes the outputs and the floor mask of V.  Faces within 2 rows of a block
// edge are computed by both neighbouring blocks (3 faces per block).
// LDS: (W*T + 3) faces x 8 x 64 doubles, then (W*T + 4) x 64 selection bytes.
template <int STAGE, int BC, int W>
__global__ __launch_bounds__(kWave * W) __attribute__((amdgpu_waves_per_eu(SWE_ADJ_WAVES_PER_EU)))
void adjoint_stage_kernel(const GridDev G, const AdjArgs A) {
    extern __shared__ double lds_faces[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const Lanes<BC> w(G, A, (int)blockIdx.x, lane);
    const int ng = G.ng, nx = G.nx;
    // row blocks of (nearly) equal height: block y takes rows [R0, Rend), at most
    // A.block_rows of them (the LDS layout)
    const int nb = (int)gridDim.y;
    const int R0 = ng + (int)(((long long)blockIdx.y * nx) / nb);
    const int Rend = ng + (int)(((long long)(blockIdx.y + 1) * nx) / nb);
    const int F0 = (R0 - 2 > ng - 1) ? R0 - 2 : ng - 1;
    const int F1 = (Rend < ng + nx - 1) ? Rend : ng + nx - 1;
    const int NF = F1 - F0 + 1;
    unsigned char *lds_sel =
        reinterpret_cast<unsigned char *>(lds_faces + (long)(A.block_rows + 3) * 8 * kWave);
    constexpr double cv = STAGE == 1 ? 1.0 : (STAGE == 2 ? 0.25 : 2.0 / 3.0);

    // ---- phase 1: x faces fa .. fb-1 of this wave
    {
        const int fa = F0 + (wv * NF) / W, fb = F0 + ((wv + 1) * NF) / W;
        if (fa < fb) {
            const double kx = G.inv_dx * (cv * *A.dt);
            bool split = G.split_ok != 0;
            bool fl;
            PCell P0, P1, P2, P3;
            D3 S0, S1, L0, L1, L2;
            int sel0, sel1;
            {
                // the four rows and two lamW rows are all in flight before the first is used
                bool gm, g0, g1, g2;
                PCell Pm = w.load_raw(fa - 1, gm);
                P0 = w.load_raw(fa, g0);
                P1 = w.load_raw(fa + 1, g1);
                P2 = w.load_raw(fa + 2, g2);
                const auto L0r = w.load_lam_raw(fa), L1r = w.load_lam_raw(fa + 1);
                w.finish(Pm, gm, fl);
                w.finish(P0, g0, fl);
                w.finish(P1, g1, fl);
                w.finish(P2, g2, fl);
                L0 = w.finish_lam(L0r);
                L1 = w.finish_lam(L1r);
                split = split && Lanes<BC>::all_lanes(
                                    Lanes<BC>::split_lanes(Pm) & Lanes<BC>::split_lanes(P0) &
                                    Lanes<BC>::split_lanes(P1) & Lanes<BC>::split_lanes(P2));
                int sa, sb, sc;
                S0.e = half_minmod(P0.eta - Pm.eta, P1.eta - P0.eta, sa);
                S0.u = half_minmod(P0.hu - Pm.hu, P1.hu - P0.hu, sb);
                S0.v = half_minmod(P0.hv - Pm.hv, P1.hv - P0.hv, sc);
                sel0 = sa | (sb << 2) | (sc << 4);
            }
#pragma unroll 1
            for (int f = fa; f < fb; ++f) {
                // row f+3 and lamW of row f+2: loaded now, finished after the face
                const int rn = f + 3 <= ng + nx + 1 ? f + 3 : ng + nx + 1;
                bool gr3;
                P3 = w.load_raw(rn, gr3);
                const auto L2r = w.load_lam_raw(f + 2);
                int sa, sb, sc;
                S1.e = half_minmod(P1.eta - P0.eta, P2.eta - P1.eta, sa);
                S1.u = half_minmod(P1.hu - P0.hu, P2.hu - P1.hu, sb);
                S1.v = half_minmod(P1.hv - P0.hv, P2.hv - P1.hv, sc);
                sel1 = sa | (sb << 2) | (sc << 4);
                // x faces: normal momentum hu, tangential hv
                FaceAdj x;
                if (split)
                    x = face_adjoint<true>(G.half_g, G.g, P0.eta + S0.e, P0.hu + S0.u,
                                           P0.hv + S0.v, P0.b, P1.eta - S1.e, P1.hu - S1.u,
                                           P1.hv - S1.v, P1.b, -kx * L0.e, -kx * L0.u,
                                           -kx * L0.v, kx * L1.e, kx * L1.u, kx * L1.v);
                else
                    x = face_adjoint<false>(G.half_g, G.g, P0.eta + S0.e, P0.hu + S0.u,
                                            P0.hv + S0.v, P0.b, P1.eta - S1.e, P1.hu - S1.u,
                                            P1.hv - S1.v, P1.b, -kx * L0.e, -kx * L0.u,
                                            -kx * L0.v, kx * L1.e, kx * L1.u, kx * L1.v);
                double *slot = lds_faces + (f - F0) * 8 * kWave + lane;
                slot[0 * kWave] = x.em;
                slot[1 * kWave] = x.nm;
                slot[2 * kWave] = x.tm;
                slot[3 * kWave] = x.bm;
                slot[4 * kWave] = x.ep;
                slot[5 * kWave] = x.np;
                slot[6 * kWave] = x.tp;
                slot[7 * kWave] = x.bp;
                lds_sel[(f - F0) * kWave + lane] = (unsigned char)sel0;
                if (f == F1) lds_sel[(f + 1 - F0) * kWave + lane] = (unsigned char)sel1;
                w.finish(P3, gr3, fl);
                L2 = w.finish_lam(L2r);
                split = split && Lanes<BC>::all_lanes(Lanes<BC>::split_lanes(P3));
                P0 = P1;
                P1 = P2;
                P2 = P3;
                S0 = S1;
                sel0 = sel1;
                L0 = L1;
                L1 = L2;
            }
        }
    }
    // ---- phase 2: rows ra .. rb-1 of this wave (an equal share of the block's rows); the
    // first row and its lamW are loaded before the barrier (ra <= Rend <= ng+nx: in the
    // padded layout even when the share is empty)
    const int ra = R0 + (wv * (Rend - R0)) / W;
    const int rb = R0 + ((wv + 1) * (Rend - R0)) / W;
    bool gra;
    PCell P = w.load_raw(ra, gra);
    const auto Lra = w.load_lam_raw(ra);
    __syncthreads();
    if (ra >= rb) return;
    const double ky = G.inv_dy * (cv * *A.dt);
    const LdsGather X{lds_faces, lds_sel, F0, lane};
    bool fl, fln;
    w.finish(P, gra, fl);
    D3 L = w.finish_lam(Lra);
    SlopeRow Sm = slope_row(slope_in(X, F1, ra - 1), G, ra - 1);
    SlopeRow S0 = slope_row(slope_in(X, F1, ra), G, ra);
#pragma unroll 1
    for (int r = ra; r < rb; ++r) {
        // the next row, its lamW, this row's read-back values and the next slope row's LDS
        // inputs are loaded first and finished after the row's arithmetic
        const int rn = r + 1 < rb ? r + 1 : r;
        bool grn;
        PCell Pn = w.load_raw(rn, grn);
        const auto Lnr = w.load_lam_raw(rn);
        const EmitIn E = load_emit<STAGE, BC>(A, w, r);
        const SlopeIn Spi = slope_in(X, F1, r + 1);
        A4 acc = (G.split_ok && Lanes<BC>::all_lanes(Lanes<BC>::split_lanes(P)))
                     ? w.template y_row<true>(G.half_g, G.g, ky, P, L)
                     : w.template y_row<false>(G.half_g, G.g, ky, P, L);
        use_emit<STAGE>(E);
        const DirectIn Di = direct_in(X, F1, r);
        const SlopeRow Sp = slope_row(Spi, G, r + 1);
        if (w.owner) {
            A.mask_in[w.row_off(r) + (w.j_bytes >> 3)] = fl ? 1 : 0;
            emit_cell<STAGE, BC>(A, w, r, L, acc,
                                 fold_ghost_rows<BC>(X, G, r, gather_rolled(G, r, Di, Sm, S0, Sp)),
                                 E);
        }
        Sm = S0;
        S0 = Sp;
        w.finish(Pn, grn, fln);
        P = Pn;
        L = w.finish_lam(Lnr);
        fl = fln;
    }
}

// Floor-state flags of a state's interior cells (the adjoint of the last step's stage 3).
__global__ __launch_bounds__(256) void floor_mask_kernel(const GridDev G, const double *h,
                                                         const double *hu, const double *hv,
                                                         unsigned char *mask) {
    const long n = (long)G.nx * G.ny;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t / G.ny), jj = (int)(t % G.ny);
        const long k = (long)(i + G.ng) * G.stride + (jj + G.ng);
        mask[k] = floor_state(h[k], hu[k], hv[k]) ? 1 : 0;
    }
}

// Zero the ghost frame (rows i < ng or i >= ng+nx, and columns j < ng or j >= ng+ny).
__global__ __launch_bounds__(256) void zero_ghosts_kernel(const GridDev G, double *a) {
    const int ng = G.ng, nx = G.nx, s = G.stride;
    const long nrow = 2L * ng * s;
    const long ncol = 2L * ng * nx;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrow + ncol) return;
    long k;
    if (t < nrow) {
        const int band = (int)(t / s);
        const int i = band < ng ? band : nx + band;
        k = (long)i * s + (t % s);
    } else {
        const long u = t - nrow;
        const int i = ng + (int)(u / (2 * ng));
        const int c = (int)(u % (2 * ng));
        k = (long)i * s + (c < ng ? c : G.ny + c);
    }
    a[k] = 0.0;
}

template <int STAGE, int BC>
hipError_t launch_adj_one(const GridDev &G, const AdjArgs &A, hipStream_t stream) {
    const dim3 block(kWave * kBlockWaves);
    const unsigned sx = (unsigned)((A.nstrips + kBlockWaves - 1) / kBlockWaves);
    hipLaunchKernelGGL((adjoint_xface_kernel<STAGE, BC>),
                       dim3(sx, (unsigned)((G.nx + 1 + A.tile_rows - 1) / A.tile_rows)), block, 0,
                       stream, G, A);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((adjoint_y_kernel<STAGE, BC>),
                       dim3(sx, (unsigned)((G.nx + A.y_rows - 1) / A.y_rows)), block, 0, stream,
                       G, A);
    return hipGetLastError();
}

// Row blocks of the fused kernel.  A block holds at most cap = W * tile_rows rows (its LDS
// layout) and keeps its slot until the slowest of its W waves is done with each phase: a
// block of h rows takes about 2 + ceil((h+3)/W) + ceil(h/W) wave-row units (the two
// prologues, its h+3 x faces, its h rows).  Blocks have floor(nx/nb) or ceil(nx/nb) rows,
// mixed evenly in dispatch order; with C resident at once, N = nstrips * nb of them take
// about floor(N/C) rounds of the mean block time plus, if a partial round is left, one
// block of the longest.  The count nb (>= ceil(nx / cap)) minimising that is taken, the
// largest of equal ones (below one round they all price as one block: more blocks fill more
// CUs).  At
// 1024^2 (18 strips, 512 resident blocks of 4 waves, cap 16) it picks 85 blocks per strip,
// 81 of them 12 rows high; pricing a block by rows + 3 alone picked 79 of 12-13 rows,
// whose 13-row blocks leave three of four waves idle in the row phase: 2% slower (84 and
// 85 against 79, 80 and 82, interleaved runs on one box; the gradient is the same bits for
// every count).
int adj_row_blocks(int nx, int nstrips, int W, int cap, long long resident) {
    auto units = [W](int h) { return 2 + (h + 3 + W - 1) / W + (h + W - 1) / W; };
    const int lo = (nx + cap - 1) / cap;
    int best = lo;
    double best_cost = -1.0;
    for (int nb = lo; nb <= nx && nb <= 4 * lo + 64; ++nb) {
        const int h = nx / nb, big = nx % nb;  // `big` blocks of h + 1 rows, the rest h
        const double mean =
            ((double)big * units(h + 1) + (double)(nb - big) * units(h)) / (double)nb;
        const long long n = (long long)nstrips * nb;
        const double cost =
            (double)(n / resident) * mean + (n % resident ? units(big ? h + 1 : h) : 0);
        if (best_cost < 0.0 || cost <= best_cost + 1e-9) {  // ties: more, shorter blocks
            best_cost = cost;
            best = nb;
        }
    }
    return best;
}

// Blocks of one fused-kernel instance resident on the device at once (CUs x occupancy),
// cached per kernel and LDS size.
long long adj_resident_blocks(const void *fn, int block_threads, size_t lds) {
    struct Key {
        const void *fn;
        size_t lds;
        int dev;
        long long n;
    };
    static thread_local std::vector<Key> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    for (const Key &k : cache)
        if (k.fn == fn && k.lds == lds && k.dev == dev) return k.n;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block_threads, lds) !=
            hipSuccess ||
        cus < 1 || per_cu < 1)
        return 256;
    const long long n = (long long)cus * per_cu;
    cache.push_back(Key{fn, lds, dev, n});
    return n;
}

template <int STAGE, int BC, int W>
hipError_t launch_fused_w(const GridDev &G, const AdjArgs &A0, hipStream_t stream) {
    AdjArgs A = A0;
    const int cap = W * A.tile_rows;
    const void *fn = reinterpret_cast<const void *>(&adjoint_stage_kernel<STAGE, BC, W>);
    const size_t lds_cap = adj_lds_bytes(W, A.tile_rows);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_cap);
    if (e != hipSuccess) return e;
    const long long resident = adj_resident_blocks(fn, kWave * W, lds_cap);
    const int nb = adj_row_blocks(G.nx, A.nstrips, W, cap, resident);
    A.block_rows = (G.nx + nb - 1) / nb;  // <= cap
    const size_t lds = adj_lds_rows_bytes(A.block_rows);
    hipLaunchKernelGGL((adjoint_stage_kernel<STAGE, BC, W>), dim3((unsigned)A.nstrips, (unsigned)nb),
                       dim3(kWave * W), lds, stream, G, A);
    return hipGetLastError();
}

template <int STAGE, int BC>
hipError_t launch_stage(const GridDev &G, const AdjArgs &A, hipStream_t stream) {
    switch (A.block_waves) {
        case 0: return launch_adj_one<STAGE, BC>(G, A, stream);
        case 2: return launch_fused_w<STAGE, BC, 2>(G, A, stream);
        case 4: return launch_fused_w<STAGE, BC, 4>(G, A, stream);
        case 8: return launch_fused_w<STAGE, BC, 8>(G, A, stream);
        default: return hipErrorInvalidValue;
    }
}

template <int BC>
hipError_t launch_adj_bc(int stage, const GridDev &G, const AdjArgs &A, hipStream_t stream) {
    switch (stage) {
        case 1: return launch_stage<1, BC>(G, A, stream);
        case 2: return launch_stage<2, BC>(G, A, stream);
        case 3: return launch_stage<3, BC>(G, A, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

int adj_strips_for(int ny) { return (ny + kAdjCols - 1) / kAdjCols; }

size_t adj_lds_rows_bytes(int block_rows) {
    const size_t faces = (size_t)block_rows + 3;
    return faces * 8 * kWave * sizeof(double) + (faces + 1) * kWave;
}

size_t adj_lds_bytes(int block_waves, int tile_rows) {
    return adj_lds_rows_bytes(block_waves * tile_rows);
}

bool adj_launch_ok(int block_waves, int tile_rows) {
    if (tile_rows < 1) return false;
    if (block_waves == 0) return true;
    if (block_waves != 2 && block_waves != 4 && block_waves != 8) return false;
    return adj_lds_bytes(block_waves, tile_rows) <= kAdjMaxLds;
}

hipError_t launch_adjoint_stage(int stage, int bc, const GridDev &G, const AdjArgs &A,
                                hipStream_t stream) {
    if (G.ng != 2 || !adj_launch_ok(A.block_waves, A.tile_rows) || A.y_rows < 1 ||
        A.nstrips != adj_strips_for(G.ny))
        return hipErrorInvalidValue;
    if (A.block_waves == 0 && (A.xf == nullptr || A.xsel == nullptr)) return hipErrorInvalidValue;
    return bc == 0 ? launch_adj_bc<0>(stage, G, A, stream) : launch_adj_bc<1>(stage, G, A, stream);
}

hipError_t launch_floor_mask(const GridDev &G, const double *h, const double *hu,
                             const double *hv, unsigned char *mask, hipStream_t stream) {
    hipLaunchKernelGGL(floor_mask_kernel, dim3(1024), dim3(256), 0, stream, G, h, hu, hv, mask);
    return hipGetLastError();
}

hipError_t launch_zero_ghosts(const GridDev &G, double *a, hipStream_t stream) {
    const long total = 2L * G.ng * G.stride + 2L * G.ng * G.nx;
    hipLaunchKernelGGL(zero_ghosts_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       stream, G, a);
    return hipGetLastError();
}

}  // namespace swe
