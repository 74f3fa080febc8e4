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
template <bool SPLIT>
__device__ __forceinline__ void velocities(double n, double t, double h, double &un,
                                           double &ut) {
    if constexpr (SPLIT) {
        const double r = div_denominator(h);
        const double qn = div_numerator(n, h, r), qt = div_numerator(t, h, r);
        un = h > kEps ? qn : 0.0;
        ut = h > kEps ? qt : 0.0;
    } else {
        un = safe_velocity(n, h);
        ut = safe_velocity(t, h);
    }
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
    double un_m, ut_m, un_p, ut_p;
    velocities<SPLIT>(n_m, t_m, h_m, un_m, ut_m);
    velocities<SPLIT>(n_p, t_p, h_p, un_p, ut_p);
    const double ym = eta_m - b_star, yp = eta_p - b_star;
    const double hL = positive_height(ym), hR = positive_height(yp);
    const double nL = hL * un_m, tL = hL * ut_m, nR = hR * un_p, tR = hR * ut_p;
    double unL, utL, unR, utR;
    velocities<SPLIT>(nL, tL, hL, unL, utL);
    velocities<SPLIT>(nR, tR, hR, unR, utR);
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
        const double r = recip(hL);
        bnL += bunL * r;
        btL += butL * r;
        bhL -= (bunL * unL + butL * utL) * r;
    }
    if (hR > kEps) {
        const double r = recip(hR);
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
        const double q = recip(h_m);
        r.nm = bun_m * q;
        r.tm = but_m * q;
        bh_m -= (bun_m * un_m + but_m * ut_m) * q;
    }
    if (h_p > kEps) {
        const double q = recip(h_p);
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

template <int STAGE, int BC>
struct AdjMarch {
    const AdjArgs &A;
    int s, ng, nx, ny, r0, r1, lane, j0;
    unsigned j_bytes, jc_bytes;
    bool gcol;   // the lane's column is outside the interior (its values are mirrored)
    bool owner;  // the lane writes its column
    bool xact;   // x faces of this column are in the domain (interior column)
    bool yact;   // the y face on the lane's upper side (j | j+1) is in the domain
    double half_g, g, kx, ky;  // kx = inv_dx * c_s * dt, ky = inv_dy * c_s * dt
    bool split;
    bool top;       // the tile starts at the first face (ng-1 | ng): it folds the top ghost rows
    double *spill;  // this wave's LDS slot for the ghost row ng-2 (4 x 64 doubles)

    __device__ __forceinline__ long long row_off(int r) const {
        return (long long)__builtin_amdgcn_readfirstlane(r * s);
    }

    __device__ __forceinline__ PCell load(int r) {
        bool gr;
        const int rs = mirror<BC>(r, ng, ng + nx, gr);
        const long long ro = row_off(rs);
        PCell c;
        const double h = ld(row_ptr(A.v_h, ro), jc_bytes);
        c.hu = ld(row_ptr(A.v_hu, ro), jc_bytes);
        c.hv = ld(row_ptr(A.v_hv, ro), jc_bytes);
        c.b = ld(row_ptr(A.B, ro), jc_bytes);
        // the floor mask of V, for the adjoint of the stage that produced V
        if (!gr && r >= r0 && r < r1 && owner)
            A.mask_in[ro + (j_bytes >> 3)] = floor_state(h, c.hu, c.hv) ? 1 : 0;
        if (BC == 0) {
            if (gr) c.hu = -c.hu;
            if (gcol) c.hv = -c.hv;
        }
        c.eta = h + c.b;
        split = split && __all(split_ok(h) && split_ok(c.hu) && split_ok(c.hv) && split_ok(c.b));
        return c;
    }

    // lamW of row r: the output adjoint where the output is not the floor state, 0 outside
    // the interior.
    __device__ __forceinline__ D3 load_lam(int r) const {
        D3 l{0.0, 0.0, 0.0};
        if (r < ng || r >= ng + nx) return l;  // wave-uniform
        const long long ro = row_off(r);
        if (xact) {
            const bool m = A.mask_out[ro + (j_bytes >> 3)] != 0;
            const double a = ld(row_ptr(A.lam_h, ro), j_bytes);
            const double b = ld(row_ptr(A.lam_hu, ro), j_bytes);
            const double c = ld(row_ptr(A.lam_hv, ro), j_bytes);
            l.e = m ? 0.0 : a;
            l.u = m ? 0.0 : b;
            l.v = m ? 0.0 : c;
        }
        return l;
    }

    // y faces of row q: adjoint contributions to the row's cells, ghost columns folded.
    template <bool SPLIT>
    __device__ __forceinline__ void y_row(const PCell &P, const D3 &L, A4 &acc) {
        const double de = from_next(P.eta) - P.eta;
        const double du = from_next(P.hu) - P.hu;
        const double dv = from_next(P.hv) - P.hv;
        int se_, su_, sv_;
        const double se = half_minmod(from_prev(de), de, se_);
        const double su = half_minmod(from_prev(du), du, su_);
        const double sv = half_minmod(from_prev(dv), dv, sv_);
        const double pe = from_next(P.eta - se);
        const double pu = from_next(P.hu - su);
        const double pv = from_next(P.hv - sv);
        const double pb = from_next(P.b);
        const double le = from_next(L.e), lu = from_next(L.u), lv = from_next(L.v);
        // y faces: normal momentum hv, tangential hu
        FaceAdj f = face_adjoint<SPLIT>(half_g, g, P.eta + se, P.hv + sv, P.hu + su, P.b, pe, pv,
                                        pu, pb, -ky * L.e, -ky * L.v, -ky * L.u, ky * le,
                                        ky * lv, ky * lu);
        if (!yact) f = face_zero();
        // plus-side parts of the face below (lane j-1) belong to this lane's cell
        const double qe = from_prev(f.ep), qn = from_prev(f.np), qt = from_prev(f.tp);
        const double qb = from_prev(f.bp);
        // slope adjoints of this cell (the minus state uses +sigma, the plus state -sigma),
        // routed by the minmod selection: own part, lane j+1's "previous" part and lane
        // j-1's "next" part land here
        double tp, ts, tn;
        slope_parts(se_, f.em - qe, tp, ts, tn);
        acc.e += f.em + qe + ts + from_next(tp) + from_prev(tn);
        slope_parts(su_, f.tm - qt, tp, ts, tn);
        acc.u += f.tm + qt + ts + from_next(tp) + from_prev(tn);
        slope_parts(sv_, f.nm - qn, tp, ts, tn);
        acc.v += f.nm + qn + ts + from_next(tp) + from_prev(tn);
        acc.b += f.bm + qb;
        fold_cols(acc);
    }

    // Ghost columns (y walls) fold into the columns they mirror, hv negated for reflecting.
    __device__ __forceinline__ void fold_cols(A4 &acc) const {
        // lane of column c is c - (j0 - 3)
        const int base = j0 - 3;
        if (j0 == ng) {  // left wall in this strip: ghosts ng-1, ng-2
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

    // One x face f | f+1 (plus the slope of row f+1 and the y faces of row f+1 when that
    // row is written by this tile).  Window on entry: P0 = row f, P1 = row f+1, L0 = lamW of
    // row f, S0/sel0 = slope of row f, Am/A0 = accumulators of rows f-1, f, pend = the
    // slope adjoint of row f collected from face f-1, P2 = row f+2, L1 = lamW of row f+1.
    // On exit: S1/sel1 = slope of row f+1, A1 = accumulator of row f+1, pend = the slope
    // adjoint of row f+1 from face f.
    template <bool SPLIT>
    __device__ __forceinline__ void step_body(int f, const PCell &P0, const PCell &P1,
                                              const PCell &P2, const D3 &L0, const D3 &L1,
                                              const D3 &S0, int sel0, D3 &S1, int &sel1, A4 &Am,
                                              A4 &A0, A4 &A1, D3 &pend) {
        int sa, sb, sc;
        S1.e = half_minmod(P1.eta - P0.eta, P2.eta - P1.eta, sa);
        S1.u = half_minmod(P1.hu - P0.hu, P2.hu - P1.hu, sb);
        S1.v = half_minmod(P1.hv - P0.hv, P2.hv - P1.hv, sc);
        sel1 = sa | (sb << 2) | (sc << 4);

        // x faces: normal momentum hu, tangential hv
        FaceAdj x = face_adjoint<SPLIT>(half_g, g, P0.eta + S0.e, P0.hu + S0.u, P0.hv + S0.v,
                                        P0.b, P1.eta - S1.e, P1.hu - S1.u, P1.hv - S1.v, P1.b,
                                        -kx * L0.e, -kx * L0.u, -kx * L0.v, kx * L1.e,
                                        kx * L1.u, kx * L1.v);
        if (!xact) x = face_zero();
        A0.e += x.em;
        A0.u += x.nm;
        A0.v += x.tm;
        A0.b += x.bm;
        A1 = A4{x.ep, x.np, x.tp, x.bp};
        // slope of row f is complete: route it to rows f-1, f, f+1
        const D3 g0{pend.e + x.em, pend.u + x.nm, pend.v + x.tm};
        pend = D3{-x.ep, -x.np, -x.tp};
        double tp, ts, tn;
        slope_parts(sel0 & 3, g0.e, tp, ts, tn);
        Am.e += tp;
        A0.e += ts;
        A1.e += tn;
        slope_parts((sel0 >> 2) & 3, g0.u, tp, ts, tn);
        Am.u += tp;
        A0.u += ts;
        A1.u += tn;
        slope_parts((sel0 >> 4) & 3, g0.v, tp, ts, tn);
        Am.v += tp;
        A0.v += ts;
        A1.v += tn;

        if (f + 1 >= r0 && f + 1 < r1) y_row<SPLIT>(P1, L1, A1);
    }

    // Loads of the next step (row f+3 of V, lamW of row f+2) are issued before this step's
    // arithmetic, so their latency hides behind one step of VALU work.
    __device__ __forceinline__ void step(int f, const PCell &P0, const PCell &P1,
                                         const PCell &P2, PCell &P3, const D3 &L0, const D3 &L1,
                                         D3 &L2, const D3 &S0, int sel0, D3 &S1, int &sel1,
                                         A4 &Am, A4 &A0, A4 &A1, D3 &pend) {
        P3 = load(f + 3 <= ng + nx + 1 ? f + 3 : ng + nx + 1);
        L2 = load_lam(f + 2);
        if (split)
            step_body<true>(f, P0, P1, P2, L0, L1, S0, sel0, S1, sel1, Am, A0, A1, pend);
        else
            step_body<false>(f, P0, P1, P2, L0, L1, S0, sel0, S1, sel1, Am, A0, A1, pend);
    }

    // Row q is complete: write it (interior rows of this tile), or fold it (ghost rows).
    __device__ __forceinline__ void retire(int q, const A4 &Aq, A4 &A0, A4 &A1) {
        if (top && q == ng - 2) {  // ghost row mirrored onto ng+1 (reflecting) / ng: not yet open
            spill[0 * kWave + lane] = Aq.e;
            spill[1 * kWave + lane] = Aq.u;
            spill[2 * kWave + lane] = Aq.v;
            spill[3 * kWave + lane] = Aq.b;
            return;
        }
        if (top && q == ng - 1) {  // rows ng (A0) and ng+1 (A1) are open
            fold_row<BC>(A0, Aq);
            const A4 g2{spill[0 * kWave + lane], spill[1 * kWave + lane],
                        spill[2 * kWave + lane], spill[3 * kWave + lane]};
            fold_row<BC>(BC == 0 ? A1 : A0, g2);
            return;
        }
        if (q < r0 || q >= r1) return;
        emit(q, Aq);
    }

    __device__ __forceinline__ void emit(int q, const A4 &Aq) const {
        if (!owner) return;
        const D3 Lq = load_lam(q);  // reloaded (a cache hit) rather than kept in registers
        const long long ro = row_off(q);
        constexpr double cv = STAGE == 1 ? 1.0 : (STAGE == 2 ? 0.25 : 2.0 / 3.0);
        double oh = cv * Lq.e + Aq.e;
        double ou = cv * Lq.u + Aq.u;
        double ov = cv * Lq.v + Aq.v;
        if constexpr (STAGE == 3) {
            st(row_ptr(A.e_h, ro), j_bytes, (1.0 / 3.0) * Lq.e);
            st(row_ptr(A.e_hu, ro), j_bytes, (1.0 / 3.0) * Lq.u);
            st(row_ptr(A.e_hv, ro), j_bytes, (1.0 / 3.0) * Lq.v);
        } else if constexpr (STAGE == 2) {
            st(row_ptr(A.e_h, ro), j_bytes, ld(row_ptr(A.e_h, ro), j_bytes) + 0.75 * Lq.e);
            st(row_ptr(A.e_hu, ro), j_bytes, ld(row_ptr(A.e_hu, ro), j_bytes) + 0.75 * Lq.u);
            st(row_ptr(A.e_hv, ro), j_bytes, ld(row_ptr(A.e_hv, ro), j_bytes) + 0.75 * Lq.v);
        } else {
            oh += ld(row_ptr(A.e_h, ro), j_bytes);
            ou += ld(row_ptr(A.e_hu, ro), j_bytes);
            ov += ld(row_ptr(A.e_hv, ro), j_bytes);
        }
        st(row_ptr(A.out_h, ro), j_bytes, oh);
        st(row_ptr(A.out_hu, ro), j_bytes, ou);
        st(row_ptr(A.out_hv, ro), j_bytes, ov);
        st(row_ptr(A.lam_b, ro), j_bytes, ld(row_ptr(A.lam_b, ro), j_bytes) + (Aq.e + Aq.b));
    }
};

#ifndef SWE_ADJ_WAVES_PER_EU
#define SWE_ADJ_WAVES_PER_EU 1
#endif

template <int STAGE, int BC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SWE_ADJ_WAVES_PER_EU)))
void adjoint_stage_kernel(const GridDev G, const AdjArgs A) {
    __shared__ double spill_lds[kBlockWaves][4 * kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int strip = blockIdx.x * kBlockWaves + wave;
    if (strip >= A.nstrips) return;

    const int ng = G.ng, nx = G.nx, ny = G.ny;
    const int j0 = ng + strip * kAdjCols;
    const int j = j0 - 3 + lane;
    const int jcl = j < ng - 2 ? ng - 2 : (j > ng + ny + 1 ? ng + ny + 1 : j);
    bool gcol;
    const int jc = mirror<BC>(jcl, ng, ng + ny, gcol);
    const int r0 = ng + (int)blockIdx.y * A.tile_rows;
    const int r1 = (r0 + A.tile_rows < ng + nx) ? r0 + A.tile_rows : ng + nx;
    const double dt = *A.dt;
    constexpr double cs = STAGE == 1 ? 1.0 : (STAGE == 2 ? 0.25 : 2.0 / 3.0);

    AdjMarch<STAGE, BC> m{A, G.stride, ng, nx, ny, r0, r1, lane, j0,
                          (unsigned)(j < 0 ? 0 : j) * 8u, (unsigned)jc * 8u,
                          j < ng || j >= ng + ny,
                          lane >= 3 && lane < 3 + kAdjCols && j < ng + ny,
                          j >= ng && j < ng + ny,
                          j >= ng - 1 && j <= ng + ny - 1,
                          G.half_g, G.g, G.inv_dx * (cs * dt), G.inv_dy * (cs * dt),
                          G.split_ok != 0, r0 - 2 <= ng - 1, &spill_lds[wave][0]};
    (void)gcol;

    const int fa = (r0 - 2 > ng - 1) ? r0 - 2 : ng - 1;
    const int fb = (r1 < ng + nx - 1) ? r1 : ng + nx - 1;
    const bool bottom = fb == ng + nx - 1;

    // prologue: rows fa-1 .. fa+2, slope of row fa, lamW of rows fa, fa+1
    PCell P0, P1, P2, P3;
    D3 S0, S1;
    int sel0, sel1;
    {
        const PCell Pm = m.load(fa - 1);
        P0 = m.load(fa);
        P1 = m.load(fa + 1);
        P2 = m.load(fa + 2);
        int sa, sb, sc;
        S0.e = half_minmod(P0.eta - Pm.eta, P1.eta - P0.eta, sa);
        S0.u = half_minmod(P0.hu - Pm.hu, P1.hu - P0.hu, sb);
        S0.v = half_minmod(P0.hv - Pm.hv, P1.hv - P0.hv, sc);
        sel0 = sa | (sb << 2) | (sc << 4);
    }
    D3 L0 = m.load_lam(fa), L1 = m.load_lam(fa + 1), L2;
    A4 Am{0.0, 0.0, 0.0, 0.0}, A0{0.0, 0.0, 0.0, 0.0}, A1;
    D3 pend{0.0, 0.0, 0.0};

#pragma unroll 1
    for (int f = fa; f <= fb; ++f) {
        m.step(f, P0, P1, P2, P3, L0, L1, L2, S0, sel0, S1, sel1, Am, A0, A1, pend);
        if (f == fb && bottom) break;
        m.retire(f - 1, Am, A0, A1);
        P0 = P1;
        P1 = P2;
        P2 = P3;
        L0 = L1;
        L1 = L2;
        S0 = S1;
        sel0 = sel1;
        Am = A0;
        A0 = A1;
    }
    if (bottom) {
        // After the last face (ng+nx-1 | ng+nx): Am = row ng+nx-2, A0 = row ng+nx-1,
        // A1 = ghost row ng+nx; the slope of ghost row ng+nx (S1/sel1, adjoint `pend`) is
        // complete and reaches ghost row ng+nx+1 too.
        A4 A2{0.0, 0.0, 0.0, 0.0};
        double tp, ts, tn;
        slope_parts(sel1 & 3, pend.e, tp, ts, tn);
        A0.e += tp;
        A1.e += ts;
        A2.e += tn;
        slope_parts((sel1 >> 2) & 3, pend.u, tp, ts, tn);
        A0.u += tp;
        A1.u += ts;
        A2.u += tn;
        slope_parts((sel1 >> 4) & 3, pend.v, tp, ts, tn);
        A0.v += tp;
        A1.v += ts;
        A2.v += tn;
        fold_row<BC>(A0, A1);                   // ng+nx   -> ng+nx-1
        fold_row<BC>(BC == 0 ? Am : A0, A2);    // ng+nx+1 -> ng+nx-2 (reflecting) / ng+nx-1
        const int q = ng + nx - 2;
        if (q >= r0 && q < r1) m.emit(q, Am);
        if (q + 1 >= r0 && q + 1 < r1) m.emit(q + 1, A0);
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
hipError_t launch_adj_one(const GridDev &G, const AdjArgs &A, hipStream_t stream, dim3 grid) {
    hipLaunchKernelGGL((adjoint_stage_kernel<STAGE, BC>), grid, dim3(kWave * kBlockWaves), 0,
                       stream, G, A);
    return hipGetLastError();
}

template <int BC>
hipError_t launch_adj_bc(int stage, const GridDev &G, const AdjArgs &A, hipStream_t stream,
                         dim3 grid) {
    switch (stage) {
        case 1: return launch_adj_one<1, BC>(G, A, stream, grid);
        case 2: return launch_adj_one<2, BC>(G, A, stream, grid);
        case 3: return launch_adj_one<3, BC>(G, A, stream, grid);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

int adj_strips_for(int ny) { return (ny + kAdjCols - 1) / kAdjCols; }

hipError_t launch_adjoint_stage(int stage, int bc, const GridDev &G, const AdjArgs &A,
                                hipStream_t stream) {
    if (G.ng != 2 || A.tile_rows < 1 || A.nstrips != adj_strips_for(G.ny))
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)((A.nstrips + kBlockWaves - 1) / kBlockWaves),
                    (unsigned)((G.nx + A.tile_rows - 1) / A.tile_rows), 1);
    return bc == 0 ? launch_adj_bc<0>(stage, G, A, stream, grid)
                   : launch_adj_bc<1>(stage, G, A, stream, grid);
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
