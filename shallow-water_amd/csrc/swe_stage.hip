// swe_stage.hip -- the fused SSP-RK3 stage kernel for gfx950 (FP64 VALU, no MFMA).
//
// One launch does, for every interior cell, what the reference does in five passes
// (include/backend/openmp/omp_solver.hpp:250-338 per stage): ghost fill, x faces, y faces,
// divergence, stage combination, positivity floor and the non-finite check; stage 3 can
// also reduce the max wave speed of its output for the next step's CFL dt.
//
// Work decomposition.  A wave (64 lanes) owns a strip of 60 consecutive y columns
// (j, contiguous in memory) plus a 2-column halo on each side, and marches down a segment
// of `tile_rows` x rows.  Along x the cross stencil (radius 2) lives in registers: each new
// row is loaded once, the per-cell minmod slope is computed once, and each x face is
// evaluated once and used by both neighbours.  Along y the neighbours live in adjacent
// lanes and are exchanged with cross-lane shuffles; each lane evaluates the y face on its
// upper side, so every y face is also computed once.  No LDS, no barriers in the march.
//
// Ghost cells are never read from memory: a row or column outside the interior is
// synthesized from the mirrored (reflecting) or nearest (transmissive) interior cell with
// the reference's sign flips (hu on x ghosts, hv on y ghosts; omp_solver.hpp:340-393).
// This equals filling the ghosts before each stage, because the stencil never touches a
// corner ghost.
//
// HBM traffic per cell and stage: read U_s (24 B) + B (8 B) [+ U^n, 24 B for stages 2, 3],
// write 24 B: 56 / 80 / 80 bytes for stages 1 / 2 / 3; halo re-reads are cache hits.
#include "swe_face.h"
#include "swe_internal.h"
#include "swe_lanes.h"

namespace swe {
namespace {

struct Cell {
    double h, hu, hv, b, eta;
};

// First column (lane 0, a halo column) of strip `strip` for ng = 2.
__device__ __forceinline__ int strip_first_col(int strip) { return strip * kStripCols; }

struct V3 {
    double e, u, v;  // (eta, hu, hv) or, for fluxes, (h, hu, hv) components
};

// State of one wave's march down its strip, and the per-row step.  The step for output row
// r sees the cells of rows r (a) and r+1 (b) and loads row r+2 into `n` first thing, using
// it only after the row's y faces (which need row r alone) have covered the load; it also
// takes the slope of row r (sa) and d_{r+1/2} = b - a (d2), and produces the slope of row
// r+1 (sb) and d_{r+3/2} (d3); fxp is the plus-side flux of face r-1/2, fxn receives that
// of face r+1/2.  The caller rotates the register names, so nothing is copied.  (Loading
// row r+3 a row earlier instead holds a fourth row in registers: 148 VGPRs against 140,
// and the step measured ~1.5% slower.)
template <int STAGE, int BC, bool SPEED, bool FAST, bool FLATB = false>
struct March {
    const StageArgs &A;
    long long base;  // element offset of this grid in the arrays (batched instances)
    int s, ng, nx, rmax, r1;
    unsigned j_bytes, jc_bytes;  // own column / source column, in bytes from the row start
    bool gcol, owner;
    double half_g, g, inv_dx, inv_dy, dt;
    double speed;
    int bad;
    // Wave-uniform: every cell of the rows finished so far by this wave satisfies split_ok,
    // so the faces of those rows may use the split IEEE division/sqrt forms (swe_face.h).
    // Sticky once false.
    bool split;

    // Row r as stored (h, hu, hv, b).  Only the loads are issued here: nothing reads the
    // values before `finish`, so the wave keeps computing while they are in flight.
    __device__ __forceinline__ Cell load_raw(int r, bool &ghost_row) {
        pin_lane_offset(jc_bytes, r);
        const int rs = mirror<BC>(r, ng, ng + nx, ghost_row);
        const long long ro = base + __builtin_amdgcn_readfirstlane(rs * s);
        Cell c;
        c.h = ld(row_ptr(A.in_h, ro), jc_bytes);
        c.hu = ld(row_ptr(A.in_hu, ro), jc_bytes);
        c.hv = ld(row_ptr(A.in_hv, ro), jc_bytes);
        // FLATB: the bottom is +0.0 everywhere, so it is not read (the same values: every
        // expression below sees b = +0.0 as it would after the load)
        c.b = FLATB ? 0.0 : ld(row_ptr(A.B, ro), jc_bytes);
        return c;
    }

    // A loaded row made usable: the reflecting ghost sign flips (momentum normal to the
    // wall negated, omp_solver.hpp:340-393), eta = h + b, and the range guard.
    __device__ __forceinline__ void finish(Cell &c, bool ghost_row) {
        if (BC == 0) {
            if (ghost_row) c.hu = -c.hu;
            if (gcol) c.hv = -c.hv;
        }
        c.eta = c.h + c.b;
        const unsigned long long exec = __builtin_amdgcn_read_exec();
        unsigned long long ok = split_ok_lanes(c.h) & split_ok_lanes(c.hu) & split_ok_lanes(c.hv);
        if (!FLATB) ok &= split_ok_lanes(c.b);
        split = split && (ok & exec) == exec;
    }

    __device__ __forceinline__ Cell load(int r) {
        bool gr;
        Cell c = load_raw(r, gr);
        finish(c, gr);
        return c;
    }

    __device__ __forceinline__ V3 load_u0(int r) {
        pin_lane_offset(jc_bytes, r);
        const long long ro = base + __builtin_amdgcn_readfirstlane(r * s);
        return V3{ld(row_ptr(A.u0_h, ro), jc_bytes), ld(row_ptr(A.u0_hu, ro), jc_bytes),
                  ld(row_ptr(A.u0_hv, ro), jc_bytes)};
    }

    // SPLIT_ONLY: the faces take the split forms without a test, and the step returns
    // false, before the x face and the stores, when the newly loaded row r+2 fails the
    // range guard (the x face of row r reads row r+2 through the slope of row r+1).
    // Otherwise each face takes the split forms while the guard holds, and the step always
    // completes (returns true).
    template <bool SPLIT_ONLY>
    __device__ __forceinline__ bool row(int r, const Cell &a, const Cell &b, Cell &n,
                                        const V3 &sa, V3 &sb, const V3 &d2, V3 &d3,
                                        const V3 &fxp, V3 &fxn) {
        // Row r+2 and U^n of row r are loaded first and not read before the y faces of row
        // r (which need row r alone, already guarded) are done: the sign flips, eta and the
        // range guard of row r+2 would otherwise wait for the loads right here.
        bool gr;
        n = load_raw(r + 2 <= rmax ? r + 2 : rmax, gr);
        V3 u0{0.0, 0.0, 0.0};
        if (STAGE >= 2) u0 = load_u0(r);
        // FAST has its own reciprocal arithmetic; only the exact mode has a split variant.
        FaceFlux fy, fx;
        if (SPLIT_ONLY || (!FAST && split))
            fy = y_faces<true>(a);
        else
            fy = y_faces<false>(a);
        finish(n, gr);
        // The unlikely marks below (guard failures, the fallback march, floor and non-finite
        // votes) keep those paths out of the hot loop's code: +0.6-0.8% on the step, the
        // same on the ensemble; marking the face bodies' own branches too moved stage 2 by
        // up to +4% on bathymetry.
        if (SPLIT_ONLY && __builtin_expect(!split, 0)) return false;
        d3 = V3{n.eta - b.eta, n.hu - b.hu, n.hv - b.hv};
        if (SPLIT_ONLY || (!FAST && split))
            fx = x_face<true>(a, b, sa, sb, d2, d3);
        else
            fx = x_face<false>(a, b, sa, sb, d2, d3);
        tail(r, a, fx, fy, fxp, fxn, u0);
        return true;
    }

    // Half the limited slope, 0.5 * minmod(a, b) (omp_kernels.hpp:146-152, :214-220): the
    // median form under the range guard (same bits), the reference form otherwise.
    template <bool SPLIT>
    static __device__ __forceinline__ double half_slope(double a, double b) {
        if constexpr (SPLIT) return half_minmod_split(a, b);
        return 0.5 * minmod(a, b);
    }

    // y faces of row r: neighbours along the lanes, face j+1/2 in lane j.  The moves fill
    // the lane without a source with 0: lane 63's face, lane 62's (it reads lane 63's slope)
    // and lane 0's slope and divergence belong to halo columns (never stored), and lanes 62
    // and 63 are left out of the face's wave votes.
    template <bool SPLIT>
    __device__ __forceinline__ FaceFlux y_faces(const Cell &a) const {
        const double dy_e = from_next_z(a.eta) - a.eta;
        const double dy_u = from_next_z(a.hu) - a.hu;
        const double dy_v = from_next_z(a.hv) - a.hv;
        const double sy_e = half_slope<SPLIT>(from_prev_z(dy_e), dy_e);
        const double sy_u = half_slope<SPLIT>(from_prev_z(dy_u), dy_u);
        const double sy_v = half_slope<SPLIT>(from_prev_z(dy_v), dy_v);
        const double pe = from_next_z(a.eta - sy_e);
        const double pu = from_next_z(a.hu - sy_u);
        const double pv = from_next_z(a.hv - sy_v);
        const double pb = FLATB ? 0.0 : from_next_z(a.b);
        // normal momentum is hv, tangential hu
        return face_flux<FAST, SPLIT>(half_g, g, a.eta + sy_e, a.hv + sy_v, a.hu + sy_u, a.b,
                                      pe, pv, pu, pb, kLastLanes);
    }

    // x face r+1/2, with the slope of row r+1 from d2 = d_{r+1/2} and d3 = d_{r+3/2}.
    template <bool SPLIT>
    __device__ __forceinline__ FaceFlux x_face(const Cell &a, const Cell &b, const V3 &sa,
                                               V3 &sb, const V3 &d2, const V3 &d3) const {
        sb = V3{half_slope<SPLIT>(d2.e, d3.e), half_slope<SPLIT>(d2.u, d3.u),
                half_slope<SPLIT>(d2.v, d3.v)};
        return face_flux<FAST, SPLIT>(half_g, g, a.eta + sa.e, a.hu + sa.u, a.hv + sa.v, a.b,
                                      b.eta - sb.e, b.hu - sb.u, b.hv - sb.v, b.b);
    }

    // Divergence of the row's four faces, the stage combination, positivity floor,
    // non-finite flag and wave speed, and the stores of row r.
    __device__ __forceinline__ void tail(int r, const Cell &a, const FaceFlux &fx,
                                         const FaceFlux &fy, const V3 &fxp, V3 &fxn,
                                         const V3 &u0) {
        fxn = V3{fx.fh, fx.fn_p, fx.ft};

        const double fyp_h = from_prev_z(fy.fh);
        const double fyp_hu = from_prev_z(fy.ft);
        const double fyp_hv = from_prev_z(fy.fn_p);

        // ---- divergence (omp_kernels.hpp:281-283)
        const double rh = divergence(inv_dx, inv_dy, fx.fh, fxp.e, fy.fh, fyp_h);
        const double rhu = divergence(inv_dx, inv_dy, fx.fn_m, fxp.u, fy.ft, fyp_hu);
        const double rhv = divergence(inv_dx, inv_dy, fx.ft, fxp.v, fy.fn_m, fyp_hv);

        const long long ro = base + __builtin_amdgcn_readfirstlane(r * s);
        if (STAGE == 0) {
            if (owner) {
                pin_lane_offset(j_bytes, r);
                st(row_ptr(A.out_h, ro), j_bytes, rh);
                st(row_ptr(A.out_hu, ro), j_bytes, rhu);
                st(row_ptr(A.out_hv, ro), j_bytes, rhv);
            }
        } else {
            // U^n is read here on every path through the row.  The combination below is used
            // only in the owner branch, which the compiler may skip when no lane owns a
            // column, and it sinks the combination into it; on that path the U^n loads would
            // still be in flight when the next row starts, and the compiler would wait for
            // them there -- and, the vector memory counter being in order, for this row's
            // stores as well -- before reusing their registers.
            if (STAGE >= 2) asm volatile("" ::"v"(u0.e), "v"(u0.u), "v"(u0.v));
            const double s0h = STAGE == 1 ? a.h : u0.e;
            const double s0u = STAGE == 1 ? a.hu : u0.u;
            const double s0v = STAGE == 1 ? a.hv : u0.v;
            double hN = rk_combine<STAGE>(s0h, a.h, dt, rh);
            double huN = rk_combine<STAGE>(s0u, a.hu, dt, rhu);
            double hvN = rk_combine<STAGE>(s0v, a.hv, dt, rhv);
            if (owner) {
                // enforce_positivity (omp_kernels.hpp:317-325); the selects run only when some
                // lane of the wave is below the floor (same values)
                if (__builtin_expect(wave_any(hN < kFloor), 0)) {
                    if (hN < kFloor) {
                        hN = kFloor;
                        huN = 0.0;
                        hvN = 0.0;
                    }
                }
                if (__builtin_expect(wave_any3(!isfinite(hN), !isfinite(huN), !isfinite(hvN)),
                                     0))
                    bad = 1;
                if (SPEED) speed = cell_speed(g, hN, huN, hvN, speed);
                pin_lane_offset(j_bytes, r);
                st(row_ptr(A.out_h, ro), j_bytes, hN);
                st(row_ptr(A.out_hu, ro), j_bytes, huN);
                st(row_ptr(A.out_hv, ro), j_bytes, hvN);
            }
        }
    }
};

// Occupancy target of the stage kernel (waves per SIMD); the VGPR budget follows from it
// (512 / waves, rounded down to the allocation granule: 168 at 3 waves).  The kernel needs
// up to 146 VGPRs and spills at 4 waves (128); on the 8192^2 dam break 3 waves measured
// ~2% faster than 4 even for an earlier, smaller version that fitted 4.  The
// persistent batch kernel has its own target.  Both are overridable at build time for
// tuning sweeps.
#ifndef SWE_STAGE_WAVES_PER_EU
#define SWE_STAGE_WAVES_PER_EU 3
#endif
#ifndef SWE_BATCH_WAVES_PER_EU
#define SWE_BATCH_WAVES_PER_EU SWE_STAGE_WAVES_PER_EU
#endif

// The march of one wave over output rows [r0, r1) from a fresh prologue (rows r0-2 .. r0+1
// and the plus-side flux of face r0-1/2).  Returns r1 when it finished, or the first row
// it did not store.
//  SPLIT_ONLY: every face takes the split division/sqrt forms (exact mode, range guard
//    held so far); the march stops at the prologue or before the x face of the first row
//    whose newly loaded row fails the guard.  The loop holds only the split face bodies,
//    so its code stays small for the instruction cache.
//  otherwise: the faces take the split forms while the guard holds and the full IEEE
//    sequences after it fails (same values either way), and the march always finishes.
// UNROLL: the loop is unrolled 6x so that the 3-row window and the 2-deep slope,
// difference and flux pairs rotate by renaming; without it they rotate by copies (the
// exact mode's fallback march, cold code).
template <bool SPLIT_ONLY, bool UNROLL, int STAGE, int BC, bool SPEED, bool FAST, bool FLATB>
__device__ __forceinline__ int march(March<STAGE, BC, SPEED, FAST, FLATB> &m, int r0, int r1) {
    // the four rows are all in flight before the first is used (one memory round trip)
    bool g2, g1, g0, gc1;
    Cell P2 = m.load_raw(r0 - 2, g2);
    Cell P1 = m.load_raw(r0 - 1, g1);
    Cell C0 = m.load_raw(r0, g0), C1 = m.load_raw(r0 + 1, gc1), C2;
    m.finish(P2, g2);
    m.finish(P1, g1);
    m.finish(C0, g0);
    m.finish(C1, gc1);
    if (SPLIT_ONLY && __builtin_expect(!m.split, 0)) return r0;

    // x differences d_{r0-3/2}, d_{r0-1/2}, d_{r0+1/2}
    const V3 dm{P1.eta - P2.eta, P1.hu - P2.hu, P1.hv - P2.hv};
    const V3 d0{C0.eta - P1.eta, C0.hu - P1.hu, C0.hv - P1.hv};
    V3 D0{C1.eta - C0.eta, C1.hu - C0.hu, C1.hv - C0.hv}, D1;
    // minus-side reconstruction of row r0-1 at face r0-1/2, slope of row r0
    const V3 sp{0.5 * minmod(dm.e, d0.e), 0.5 * minmod(dm.u, d0.u), 0.5 * minmod(dm.v, d0.v)};
    V3 S0{0.5 * minmod(d0.e, D0.e), 0.5 * minmod(d0.u, D0.u), 0.5 * minmod(d0.v, D0.v)}, S1;
    const FaceFlux f0 =
        (SPLIT_ONLY || (!FAST && m.split))
            ? face_flux<FAST, true>(m.half_g, m.g, P1.eta + sp.e, P1.hu + sp.u, P1.hv + sp.v,
                                    P1.b, C0.eta - S0.e, C0.hu - S0.u, C0.hv - S0.v, C0.b)
            : face_flux<FAST, false>(m.half_g, m.g, P1.eta + sp.e, P1.hu + sp.u, P1.hv + sp.v,
                                     P1.b, C0.eta - S0.e, C0.hu - S0.u, C0.hv - S0.v, C0.b);
    V3 F0{f0.fh, f0.fn_p, f0.ft}, F1;

    if constexpr (UNROLL) {
        for (int r = r0; r < r1; r += 6) {
            if (!m.template row<SPLIT_ONLY>(r, C0, C1, C2, S0, S1, D0, D1, F0, F1)) return r;
            if (r + 1 >= r1) break;
            if (!m.template row<SPLIT_ONLY>(r + 1, C1, C2, C0, S1, S0, D1, D0, F1, F0))
                return r + 1;
            if (r + 2 >= r1) break;
            if (!m.template row<SPLIT_ONLY>(r + 2, C2, C0, C1, S0, S1, D0, D1, F0, F1))
                return r + 2;
            if (r + 3 >= r1) break;
            if (!m.template row<SPLIT_ONLY>(r + 3, C0, C1, C2, S1, S0, D1, D0, F1, F0))
                return r + 3;
            if (r + 4 >= r1) break;
            if (!m.template row<SPLIT_ONLY>(r + 4, C1, C2, C0, S0, S1, D0, D1, F0, F1))
                return r + 4;
            if (r + 5 >= r1) break;
            if (!m.template row<SPLIT_ONLY>(r + 5, C2, C0, C1, S1, S0, D1, D0, F1, F0))
                return r + 5;
        }
    } else {
#pragma clang loop unroll(disable)
        for (int r = r0; r < r1; ++r) {
            if (!m.template row<SPLIT_ONLY>(r, C0, C1, C2, S0, S1, D0, D1, F0, F1)) return r;
            C0 = C1;
            C1 = C2;
            S0 = S1;
            D0 = D1;
            F0 = F1;
        }
    }
    return r1;
}

// One wave's march over output rows [r0, r1) (interior rows, r0 < r1) with lane l on
// column col0 + l, writing the columns of lanes [own_lo, own_hi) that are interior.  Input
// columns outside the 2-cell ghost frame are clamped to its edge: such lanes feed only
// lanes outside [own_lo, own_hi) (the caller's halo), never a stored value.  Accumulates
// the max wave speed of the written cells (SPEED) and a non-finite flag into the caller's
// per-lane values.
//
// In the exact mode the split-only march runs first; if the range guard fails in a row,
// the fallback march restarts from that row with its own prologue (the rows before it are
// stored, and the faces it recomputes have the same values in either form).
template <int STAGE, int BC, bool SPEED, bool FAST, bool FLATB = false>
__device__ __forceinline__ void run_rows(const GridDev &G, const StageArgs &A, long long base,
                                         double dt, int col0, int own_lo, int own_hi, int r0,
                                         int r1, int lane, double &speed, int &bad) {
    const int ng = G.ng, nx = G.nx, ny = G.ny;
    const int j = col0 + lane;
    bool gcol;
    const int jl = j < ng - 2 ? ng - 2 : (j < ng + ny + 2 ? j : ng + ny + 1);
    const int jc = mirror<BC>(jl, ng, ng + ny, gcol);

    March<STAGE, BC, SPEED, FAST, FLATB> m{A, base, G.stride, ng, nx, ng + nx + 1, r1,
                                    (unsigned)jl * 8u, (unsigned)jc * 8u, gcol,
                                    lane >= own_lo && lane < own_hi && j >= ng && j < ng + ny,
                                    G.half_g, G.g, G.inv_dx, G.inv_dy, dt, speed, bad,
                                    G.split_ok != 0};
    int r = r0;
    if (!FAST && G.split_ok) r = march<true, true>(m, r0, r1);
    if (__builtin_expect(r < r1, 0)) {
        m.split = G.split_ok != 0;
        march<false, true>(m, r, r1);
    }
    speed = m.speed;
    bad = m.bad;
}

// One wave's tile of a stage launch: strip `strip` (60 output columns) of rows
// [r0, min(r0 + tile_rows, ng + nx)).
template <int STAGE, int BC, bool SPEED, bool FAST, bool FLATB = false>
__device__ __forceinline__ void run_tile(const GridDev &G, const StageArgs &A, long long base,
                                         double dt, int strip, int r0, int lane, double &speed,
                                         int &bad) {
    const int r1 = (r0 + A.tile_rows < G.ng + G.nx) ? r0 + A.tile_rows : G.ng + G.nx;
    run_rows<STAGE, BC, SPEED, FAST, FLATB>(G, A, base, dt, G.ng - 2 + strip * kStripCols, 2,
                                     kWave - 2, r0, r1, lane, speed, bad);
}

// Workgroup -> tile order.  The dispatcher hands consecutive workgroup ids of a 1-D grid
// to the XCDs in turn (id mod 8), and each XCD has its own L2.  xcd_order gives each XCD
// runs of `chunk` consecutive tiles: workgroups b = x + 8k (XCD x) take tiles
// G*8*chunk + x*chunk + k%chunk with G = k / chunk, so the tiles that run side by side on
// one XCD are neighbours in the tile order (horizontal neighbours whose halo columns share
// cache lines in that XCD's L2), while successive runs still rotate over the XCDs.  The
// workgroups past the last whole round of 8*chunk keep their own index.  A bijection of
// [0, total).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_order(int b, int total, int chunk) {
    const int span = kXcds * chunk;
    if (chunk <= 1 || b >= (total / span) * span) return b;
    const int x = b % kXcds, k = b / kXcds;
    return (k / chunk) * span + x * chunk + k % chunk;
}
// The stage kernel's tiles all cost the same, so each XCD takes one eighth of the grid.
#ifndef SWE_STAGE_XCD_REMAP
#define SWE_STAGE_XCD_REMAP 1
#endif
__device__ __forceinline__ int xcd_tile(int b, int total) {
    return SWE_STAGE_XCD_REMAP ? xcd_order(b, total, total / kXcds) : b;
}
// Batch tiles differ in length (full segments first, short last segments at the end of
// the work-list): short runs keep the XCDs' shares of long and short tiles even.
#ifndef SWE_BATCH_XCD_CHUNK
#define SWE_BATCH_XCD_CHUNK 16
#endif

template <int STAGE, int BC, bool SPEED, bool FAST, bool FLATB>
__global__ __launch_bounds__(kWave * kStageBlockWaves)
__attribute__((amdgpu_waves_per_eu(SWE_STAGE_WAVES_PER_EU)))
void stage_kernel(const GridDev G, const StageArgs A) {
    // scalar loads issued first, ahead of the launch's first store: dt and the other
    // stages' run status words (kStatusWords)
    const double dt = STAGE != 0 ? *A.dt : 0.0;
    const int stop = STAGE != 0 ? stage_stop_words<STAGE>(A.flag) : 0;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    // 1-D grid of (row segment, strip group) tiles, segment-major
    const int groups = (A.nstrips + kStageBlockWaves - 1) / kStageBlockWaves;
    const int t = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int seg = t / groups;
    const int strip = (t - seg * groups) * kStageBlockWaves + wave;

    double speed = 0.0;
    int bad = 0;

    if constexpr (STAGE != 0) {
        if (stop) {  // an earlier stage of this run went non-finite
            stage_skip<STAGE>(A.flag);
            return;
        }
        if (STAGE == 1 && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(A.flag, 1);
    }

    if (strip < A.nstrips)
        run_tile<STAGE, BC, SPEED, FAST, FLATB>(G, A, 0, dt, strip, G.ng + seg * A.tile_rows,
                                                lane, speed, bad);

    if (STAGE != 0) {
        if (__any(bad)) {
            if (lane == 0) atomicOr(A.flag + STAGE, kStatusBad);
        }
    }
    if (SPEED) {
        for (int off = kWave / 2; off > 0; off >>= 1) speed = smax(speed, __shfl_xor(speed, off));
        if constexpr (kStageBlockWaves == 1) {
            if (lane == 0) A.partial[blockIdx.x] = speed;
        } else {
            __shared__ double red[kStageBlockWaves];
            if (lane == 0) red[wave] = speed;
            __syncthreads();
            if (threadIdx.x == 0) {
                double mx = red[0];
                for (int w = 1; w < kStageBlockWaves; ++w) mx = smax(mx, red[w]);
                A.partial[blockIdx.x] = mx;
            }
        }
    }
}

// ---- one launch per step: the three stages fused (temporal blocking) -------------------
//
// A wave owns kFusedCols = 52 output columns and a segment of rows [r0, r1), and runs the
// three stages of the step over it one after the other, each with the march above:
//   stage 1 over rows [r0-4, r1+4) -> U1, writing lanes [2, 62)  (columns c0-4 .. c0+55)
//   stage 2 over rows [r0-2, r1+2) -> U2, writing lanes [4, 60)  (columns c0-2 .. c0+53)
//   stage 3 over rows [r0, r1)      -> U^{n+1}, writing lanes [6, 58) (columns c0 .. c0+51)
// (rows clipped to the interior; lane l on column c0 - 6 + l).  Each stage widens the
// stencil by 2 cells on every side, so every stage-2 and stage-3 value a wave reads is one
// it wrote itself in the stage before: the ghost columns and rows it reads are mirrored
// loads of interior cells it wrote, and the lanes beyond that only feed its own halo
// lanes.  Neighbouring tiles overlap and write the same cells of U1 and U2 -- with the
// same bits, as every cell's arithmetic is the same in every tile -- so no tile depends on
// another; a workgroup-scope fence orders each wave's stores before its own loads of the
// next stage (other lanes' cells, through the mirrored loads).  U^{n+1} goes to a separate
// array (U^n is still read by other tiles' halos), and the caller swaps the two.
//
// Per output cell this computes 64/52 lanes x (3 + 12/(r1-r0)) row-stages of face
// arithmetic against 64/60 x 3 for the three launches; the HBM traffic falls to U^n and B
// read once, U^{n+1} written once, plus U1 and U2 written (and read back from cache).
//
// Run status (kStatusWords): launch k of a call writes word 1 + (k & 1) -- bit s-1 when
// stage s produced a non-finite value, kFusedSkipped when it did nothing -- and reads the
// other word, written by launch k-1: one set, and it does nothing but mark its own word.
// A launch never reads the word it writes, so the failing launch runs to its end.
constexpr int kFusedCols = kWave - 12;
constexpr int kFusedSkipped = 8;

template <int BC, bool SPEED, bool FAST, bool FLATB>
__global__ __launch_bounds__(kWave * kStageBlockWaves)
__attribute__((amdgpu_waves_per_eu(SWE_STAGE_WAVES_PER_EU)))
void fused_step_kernel(const GridDev G, const StageArgs A1, const StageArgs A2,
                       const StageArgs A3, int parity) {
    const double dt = *A1.dt;
    int *const status = A1.flag;
    const int stop = status[2 - parity];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int groups = (A1.nstrips + kStageBlockWaves - 1) / kStageBlockWaves;
    const int t = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int seg = t / groups;
    const int strip = (t - seg * groups) * kStageBlockWaves + wave;
    if (stop) {  // an earlier step of this call went non-finite
        if (threadIdx.x == 0) atomicOr(status + 1 + parity, kFusedSkipped);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(status, 1);

    double speed = 0.0;
    int bad1 = 0, bad2 = 0, bad3 = 0;
    if (strip < A1.nstrips) {
        const int ng = G.ng, ie = G.ng + G.nx;
        const int col0 = ng - 6 + strip * kFusedCols;
        const int r0 = ng + seg * A1.tile_rows;
        const int r1 = (r0 + A1.tile_rows < ie) ? r0 + A1.tile_rows : ie;
        double unused = 0.0;
        run_rows<1, BC, false, FAST, FLATB>(G, A1, 0, dt, col0, 2, kWave - 2,
                                            r0 - 4 > ng ? r0 - 4 : ng,
                                            r1 + 4 < ie ? r1 + 4 : ie, lane, unused, bad1);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        run_rows<2, BC, false, FAST, FLATB>(G, A2, 0, dt, col0, 4, kWave - 4,
                                            r0 - 2 > ng ? r0 - 2 : ng,
                                            r1 + 2 < ie ? r1 + 2 : ie, lane, unused, bad2);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        run_rows<3, BC, SPEED, FAST, FLATB>(G, A3, 0, dt, col0, 6, kWave - 6, r0, r1, lane,
                                            speed, bad3);
    }
    const int bits = (__any(bad1) ? 1 : 0) | (__any(bad2) ? 2 : 0) | (__any(bad3) ? 4 : 0);
    if (bits != 0 && lane == 0) atomicOr(status + 1 + parity, bits);
    if (SPEED) {
        for (int off = kWave / 2; off > 0; off >>= 1) speed = smax(speed, __shfl_xor(speed, off));
        __shared__ double red[kStageBlockWaves];
        if (lane == 0) red[wave] = speed;
        __syncthreads();
        if (threadIdx.x == 0) {
            double mx = red[0];
            for (int w = 1; w < kStageBlockWaves; ++w) mx = smax(mx, red[w]);
            A3.partial[blockIdx.x] = mx;
        }
    }
}

// ---- batched ensemble ----------------------------------------------------------------

__device__ __forceinline__ GridDev inst_grid(const BatchArgs &B, const BatchInst &I) {
    return GridDev{I.nx, I.ny, 2, I.stride, I.dx, I.dy, I.inv_dx, I.inv_dy, B.g, B.half_g,
                   B.split_ok};
}

__device__ __forceinline__ double wave_max(double v) {
    for (int off = kWave / 2; off > 0; off >>= 1) v = smax(v, __shfl_xor(v, off));
    return v;
}

// Persistent tile loop.  Every wave walks the work-list with stride W; the tile index is
// wave-uniform, so the tile and instance records are scalar loads.  Per-instance max wave
// speed goes to max_bits[inst] with a 64-bit atomic max on the bit pattern (exact for
// non-negative doubles, order-independent); non-finite flags per instance.
template <int STAGE, int BC, bool SPEED, bool FAST>
__global__ __launch_bounds__(kWave * kBatchBlockWaves)
__attribute__((amdgpu_waves_per_eu(SWE_BATCH_WAVES_PER_EU)))
void batch_stage_kernel(const BatchArgs B) {
    const int lane = threadIdx.x & (kWave - 1);
    // workgroups in XCD-aware order (xcd_tile), so neighbouring tiles of the work-list run
    // on the same XCD
    const int gw = __builtin_amdgcn_readfirstlane(
        xcd_order((int)blockIdx.x, (int)gridDim.x, SWE_BATCH_XCD_CHUNK) * kBatchBlockWaves +
        (int)(threadIdx.x / kWave));
    const int nw = (int)(gridDim.x * blockDim.x / kWave);
    for (int t = gw; t < B.ntiles; t += nw) {
        const BatchTile T = B.tiles[t];
        const BatchInst I = B.inst[T.inst];
        int *const status = B.flags + (size_t)kStatusWords * T.inst;
        if constexpr (STAGE != 0) {
            // the instance's first-bad-stage stop (kStatusWords): the words of the two other
            // stages, which no wave of this launch writes
            constexpr int a = STAGE % 3 + 1, b = (STAGE + 1) % 3 + 1;
            if (__builtin_amdgcn_readfirstlane(status[a] | status[b]) != 0) {
                if (lane == 0) atomicOr(status + STAGE, kStatusSkipped);
                continue;
            }
            if (STAGE == 1 && T.strip == 0 && T.r0 == 2 && lane == 0) atomicAdd(status, 1);
        }
        const GridDev G = inst_grid(B, I);
        double speed = 0.0;
        int bad = 0;
        run_tile<STAGE, BC, SPEED, FAST>(G, B.base, I.offset, STAGE != 0 ? B.dt[T.inst] : 0.0,
                                         T.strip, T.r0, lane, speed, bad);
        if (SPEED) {
            speed = wave_max(speed);
            if (lane == 0) atomicMax(&B.max_bits[T.inst], (unsigned long long)__double_as_longlong(speed));
        }
        if (STAGE != 0 && __any(bad)) {
            if (lane == 0) atomicOr(status + STAGE, kStatusBad);
        }
    }
}

// Max wave speed of U over each tile's cells (the owner lanes of its rows).
__global__ __launch_bounds__(256) void batch_speed_kernel(const BatchArgs B) {
    const int lane = threadIdx.x & (kWave - 1);
    const int gw = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave));
    const int nw = (int)(gridDim.x * blockDim.x / kWave);
    for (int t = gw; t < B.ntiles; t += nw) {
        const BatchTile T = B.tiles[t];
        const BatchInst I = B.inst[T.inst];
        const int j = strip_first_col(T.strip) + lane;
        const bool owner = lane >= 2 && lane < kWave - 2 && j < 2 + I.ny;
        const int r1 = (T.r0 + B.base.tile_rows < 2 + I.nx) ? T.r0 + B.base.tile_rows : 2 + I.nx;
        double m = 0.0;
        if (owner) {
            const double *h = B.base.in_h + I.offset, *hu = B.base.in_hu + I.offset,
                         *hv = B.base.in_hv + I.offset;
            for (int r = T.r0; r < r1; ++r) {
                const long long k = (long long)r * I.stride + j;
                m = cell_speed(B.g, h[k], hu[k], hv[k], m);
            }
        }
        m = wave_max(m);
        if (lane == 0) atomicMax(&B.max_bits[T.inst], (unsigned long long)__double_as_longlong(m));
    }
}

__global__ __launch_bounds__(1024) void batch_finalize_kernel(const BatchInst *inst, int count,
                                                             unsigned long long *max_bits,
                                                             double cfl, double *dt,
                                                             double *hist, long long *counter,
                                                             long long hist_rows,
                                                             const int *flags) {
    const long long c = counter ? *counter : 0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        const int *st = flags != nullptr ? flags + (size_t)kStatusWords * i : nullptr;
        const bool stopped = st != nullptr && (st[1] | st[2] | st[3]) != 0;
        const double ms = __longlong_as_double((long long)max_bits[i]);
        const double d = stopped ? __builtin_nan("") : dt_from_speed(ms, cfl, inst[i].dx, inst[i].dy);
        if (!stopped) dt[i] = d;
        max_bits[i] = 0ull;
        if (hist != nullptr && c < hist_rows) hist[c * count + i] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0 && counter != nullptr) *counter = c + 1;
}

template <int STAGE, int BC, bool SPEED, bool FAST>
hipError_t launch_batch_one(const BatchArgs &B, int blocks, hipStream_t stream) {
    hipLaunchKernelGGL((batch_stage_kernel<STAGE, BC, SPEED, FAST>), dim3(blocks),
                       dim3(kWave * kBatchBlockWaves), 0, stream, B);
    return hipGetLastError();
}

template <int BC, bool FAST>
hipError_t launch_batch_bc(int stage, bool speed, const BatchArgs &B, int blocks,
                           hipStream_t stream) {
    switch (stage) {
        case 1: return launch_batch_one<1, BC, false, FAST>(B, blocks, stream);
        case 2: return launch_batch_one<2, BC, false, FAST>(B, blocks, stream);
        case 3:
            return speed ? launch_batch_one<3, BC, true, FAST>(B, blocks, stream)
                         : launch_batch_one<3, BC, false, FAST>(B, blocks, stream);
        default: return hipErrorInvalidValue;
    }
}

__global__ __launch_bounds__(256) void speed_reduce_kernel(const GridDev G, const double *h,
                                                           const double *hu, const double *hv,
                                                           double *partial) {
    const long n = (long)G.nx * G.ny;
    double m = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t / G.ny), jj = (int)(t % G.ny);
        const long k = (long)(i + G.ng) * G.stride + (jj + G.ng);
        m = cell_speed(G.g, h[k], hu[k], hv[k], m);
    }
    for (int off = kWave / 2; off > 0; off >>= 1) m = smax(m, __shfl_xor(m, off));
    __shared__ double red[256 / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    if (lane == 0) red[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < (int)(blockDim.x / kWave); ++w) r = smax(r, red[w]);
        partial[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(256) void finalize_dt_kernel(const double *partial, int n, double cfl,
                                                          double dx, double dy, double *scal,
                                                          double *hist, long long *counter,
                                                          long long hist_cap,
                                                          const int *status) {
    double m = 0.0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) m = smax(m, partial[t]);
    for (int off = kWave / 2; off > 0; off >>= 1) m = smax(m, __shfl_xor(m, off));
    __shared__ double red[256 / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    if (lane == 0) red[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < (int)(blockDim.x / kWave); ++w) r = smax(r, red[w]);
        const bool stopped = status != nullptr && (status[1] | status[2] | status[3]) != 0;
        const double dt = stopped ? __builtin_nan("") : dt_from_speed(r, cfl, dx, dy);
        if (!stopped) {
            scal[0] = dt;
            scal[1] = r;
        }
        if (hist != nullptr) {
            const long long c = *counter;
            if (c < hist_cap) hist[c] = dt;
            *counter = c + 1;
        }
    }
}

// One thread per ghost cell.  Threads [0, 2*ng*nyt) cover the full ghost rows
// (i < ng or i >= ng+nx, every j); the rest cover ghost columns of interior rows.
template <int BC>
__global__ __launch_bounds__(256) void ghost_fill_kernel(const GridDev G, double *h, double *hu,
                                                         double *hv) {
    const int ng = G.ng, nx = G.nx, ny = G.ny, s = G.stride;
    const long nrow = 2L * ng * s;
    const long ncol = 2L * ng * nx;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrow + ncol) return;
    int i, j;
    if (t < nrow) {
        const int band = (int)(t / s);  // 0..2ng-1
        j = (int)(t % s);
        i = band < ng ? band : nx + band;  // ng..2ng-1 -> ng+nx .. 2ng+nx-1
    } else {
        const long u = t - nrow;
        i = ng + (int)(u / (2 * ng));
        const int c = (int)(u % (2 * ng));
        j = c < ng ? c : ny + c;
    }
    bool gi, gj;
    const int si = mirror<BC>(i, ng, ng + nx, gi);
    const int sj = mirror<BC>(j, ng, ng + ny, gj);
    // x pass (omp_solver.hpp:346-368) fills ghost rows at interior j from the mirrored row
    // with hu negated; the y pass (:370-392) then fills every i's ghost columns from the
    // mirrored column of the (possibly x-filled) row with hv negated.
    const long k = (long)i * s + j;
    const long src = (long)si * s + sj;
    h[k] = h[src];
    if (hu != nullptr) {
        double vu = hu[src], vv = hv[src];
        if (BC == 0) {
            if (gi) vu = -vu;
            if (gj) vv = -vv;
        }
        hu[k] = vu;
        hv[k] = vv;
    }
}

__global__ void set_scalar_kernel(double *dst, double v) { *dst = v; }

__global__ void fill_int_kernel(int *dst, int v, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = v;
}

__global__ __launch_bounds__(256) void nonzero_bits_kernel(const double *a, long n,
                                                           int *flag) {
    bool nz = false;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x)
        nz = nz || __double_as_longlong(a[t]) != 0;
    if (wave_any(nz) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(flag, 1);
}

__global__ void fill_double_kernel(double *dst, double v, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = v;
}

template <int STAGE, int BC, bool SPEED, bool FAST, bool FLATB>
hipError_t launch_one(const GridDev &G, const StageArgs &A, hipStream_t stream, dim3 grid) {
    hipLaunchKernelGGL((stage_kernel<STAGE, BC, SPEED, FAST, FLATB>), grid,
                       dim3(kWave * kStageBlockWaves), 0, stream, G, A);
    return hipGetLastError();
}

template <int BC, bool FAST, bool FLATB = false>
hipError_t launch_bc(int stage, bool speed, const GridDev &G, const StageArgs &A,
                     hipStream_t stream, dim3 grid) {
    switch (stage) {
        case 0: return launch_one<0, BC, false, FAST, FLATB>(G, A, stream, grid);
        case 1: return launch_one<1, BC, false, FAST, FLATB>(G, A, stream, grid);
        case 2: return launch_one<2, BC, false, FAST, FLATB>(G, A, stream, grid);
        case 3:
            return speed ? launch_one<3, BC, true, FAST, FLATB>(G, A, stream, grid)
                         : launch_one<3, BC, false, FAST, FLATB>(G, A, stream, grid);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

dim3 stage_grid(const GridDev &G, int tile_rows, int block_waves) {
    const int nstrips = strips_for(G.ny);
    return dim3((nstrips + block_waves - 1) / block_waves, (G.nx + tile_rows - 1) / tile_rows, 1);
}

hipError_t launch_stage(int stage, int bc, bool speed, bool fast, const GridDev &G,
                        const StageArgs &A, hipStream_t stream, dim3 *grid_out) {
    if (A.tile_rows < 1 || A.nstrips != strips_for(G.ny)) return hipErrorInvalidValue;
    const dim3 g2 = stage_grid(G, A.tile_rows);
    const dim3 grid(g2.x * g2.y, 1, 1);  // segment-major tiles, see xcd_tile
    if (grid_out) *grid_out = grid;
    if (fast)
        return bc == 0 ? launch_bc<0, true>(stage, speed, G, A, stream, grid)
                       : launch_bc<1, true>(stage, speed, G, A, stream, grid);
    // a bottom that is +0.0 everywhere: the exact-mode kernels that do not read it
    if (A.flat_b)
        return bc == 0 ? launch_bc<0, false, true>(stage, speed, G, A, stream, grid)
                       : launch_bc<1, false, true>(stage, speed, G, A, stream, grid);
    return bc == 0 ? launch_bc<0, false>(stage, speed, G, A, stream, grid)
                   : launch_bc<1, false>(stage, speed, G, A, stream, grid);
}

int fused_strips_for(int ny) { return (ny + kFusedCols - 1) / kFusedCols; }

int fused_step_blocks(const GridDev &G, int tile_rows) {
    const long long groups = (fused_strips_for(G.ny) + kStageBlockWaves - 1) / kStageBlockWaves;
    return (int)(groups * ((G.nx + tile_rows - 1) / tile_rows));
}

hipError_t launch_fused_step(int bc, bool speed, bool fast, const GridDev &G,
                             const StageArgs &A1, const StageArgs &A2, const StageArgs &A3,
                             int parity, hipStream_t stream, dim3 *grid_out) {
    if (G.ng != 2 || A1.tile_rows < 4 || A1.nstrips != fused_strips_for(G.ny) ||
        (parity & ~1) != 0 || A3.out_h == A3.u0_h || A3.out_h == A3.in_h)
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)fused_step_blocks(G, A1.tile_rows), 1, 1);
    if (grid_out) *grid_out = grid;
    const dim3 block(kWave * kStageBlockWaves);
#define SWE_FUSED(BCV, SP, FA, FB)                                                   \
    hipLaunchKernelGGL((fused_step_kernel<BCV, SP, FA, FB>), grid, block, 0, stream, G, A1, \
                       A2, A3, parity)
#define SWE_FUSED_SP(BCV, FA, FB)            \
    do {                                     \
        if (speed)                           \
            SWE_FUSED(BCV, true, FA, FB);    \
        else                                 \
            SWE_FUSED(BCV, false, FA, FB);   \
    } while (0)
    // the exact mode with a bottom of +0.0 everywhere: the marches that do not read B
    const bool fb = !fast && A1.flat_b;
    if (bc == 0) {
        if (fast) SWE_FUSED_SP(0, true, false);
        else if (fb) SWE_FUSED_SP(0, false, true);
        else SWE_FUSED_SP(0, false, false);
    } else {
        if (fast) SWE_FUSED_SP(1, true, false);
        else if (fb) SWE_FUSED_SP(1, false, true);
        else SWE_FUSED_SP(1, false, false);
    }
#undef SWE_FUSED_SP
#undef SWE_FUSED
    return hipGetLastError();
}

hipError_t launch_batch_stage(int stage, int bc, bool speed, bool fast, const BatchArgs &B,
                              int blocks, hipStream_t stream) {
    if (blocks < 1 || B.ntiles < 1 || B.base.tile_rows < 1) return hipErrorInvalidValue;
    if (fast)
        return bc == 0 ? launch_batch_bc<0, true>(stage, speed, B, blocks, stream)
                       : launch_batch_bc<1, true>(stage, speed, B, blocks, stream);
    return bc == 0 ? launch_batch_bc<0, false>(stage, speed, B, blocks, stream)
                   : launch_batch_bc<1, false>(stage, speed, B, blocks, stream);
}

hipError_t launch_batch_speed(const BatchArgs &B, int blocks, hipStream_t stream) {
    hipLaunchKernelGGL(batch_speed_kernel, dim3(blocks), dim3(kWave * kBatchBlockWaves), 0,
                       stream, B);
    return hipGetLastError();
}

hipError_t launch_batch_finalize(const BatchInst *inst, int count, unsigned long long *max_bits,
                                 double cfl, double *dt, double *hist, long long *counter,
                                 long long hist_rows, const int *flags, hipStream_t stream) {
    hipLaunchKernelGGL(batch_finalize_kernel, dim3(1), dim3(1024), 0, stream, inst, count,
                       max_bits, cfl, dt, hist, counter, hist_rows, flags);
    return hipGetLastError();
}

hipError_t launch_speed_reduce(const GridDev &G, const double *h, const double *hu,
                               const double *hv, double *partial, int blocks,
                               hipStream_t stream) {
    hipLaunchKernelGGL(speed_reduce_kernel, dim3(blocks), dim3(256), 0, stream, G, h, hu, hv,
                       partial);
    return hipGetLastError();
}

hipError_t launch_finalize_dt(const double *partial, int n, double cfl, double dx, double dy,
                              double *scal, double *hist, long long *counter, long long hist_cap,
                              const int *status, hipStream_t stream) {
    hipLaunchKernelGGL(finalize_dt_kernel, dim3(1), dim3(256), 0, stream, partial, n, cfl, dx, dy,
                       scal, hist, counter, hist_cap, status);
    return hipGetLastError();
}

hipError_t launch_ghost_fill(int bc, const GridDev &G, double *h, double *hu, double *hv,
                             hipStream_t stream) {
    const long total = 2L * G.ng * G.stride + 2L * G.ng * G.nx;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (bc == 0)
        hipLaunchKernelGGL(ghost_fill_kernel<0>, grid, dim3(256), 0, stream, G, h, hu, hv);
    else
        hipLaunchKernelGGL(ghost_fill_kernel<1>, grid, dim3(256), 0, stream, G, h, hu, hv);
    return hipGetLastError();
}

hipError_t launch_set_scalar(double *dst, double v, hipStream_t stream) {
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, stream, dst, v);
    return hipGetLastError();
}

hipError_t launch_fill_int(int *dst, int v, int n, hipStream_t stream) {
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, dst, v, n);
    return hipGetLastError();
}

hipError_t launch_nonzero_bits(const double *a, long n, int *flag, hipStream_t stream) {
    hipLaunchKernelGGL(nonzero_bits_kernel, dim3(1024), dim3(256), 0, stream, a, n, flag);
    return hipGetLastError();
}

hipError_t launch_fill_double(double *dst, double v, long n, hipStream_t stream) {
    hipLaunchKernelGGL(fill_double_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       dst, v, n);
    return hipGetLastError();
}

}  // namespace swe
