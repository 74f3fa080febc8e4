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

namespace swe {
namespace {

struct Cell {
    double h, hu, hv, b, eta;
};

// Lane l receives lane l+1's value / lane l-1's value.  Only lanes 1..62 consume them.
__device__ __forceinline__ double from_next(double x) { return __shfl_down(x, 1); }
__device__ __forceinline__ double from_prev(double x) { return __shfl_up(x, 1); }

template <int BC>
__device__ __forceinline__ int mirror(int r, int lo, int hi, bool &ghost) {
    // lo..hi-1 is the interior range; r may lie up to 2 cells outside it.
    ghost = false;
    if (r < lo) {
        ghost = true;
        return BC == 0 ? 2 * lo - 1 - r : lo;
    }
    if (r >= hi) {
        ghost = true;
        return BC == 0 ? 2 * hi - 1 - r : hi - 1;
    }
    return r;
}

template <int BC>
__device__ __forceinline__ Cell load_cell(const StageArgs &A, int row_off, bool ghost_row,
                                          int jc, bool ghost_col) {
    const int k = row_off + jc;
    Cell c;
    c.h = A.in_h[k];
    c.hu = A.in_hu[k];
    c.hv = A.in_hv[k];
    c.b = A.B[k];
    if (BC == 0) {
        if (ghost_row) c.hu = -c.hu;
        if (ghost_col) c.hv = -c.hv;
    }
    c.eta = c.h + c.b;
    return c;
}

template <int STAGE, int BC, bool SPEED>
__global__ __launch_bounds__(256) void stage_kernel(const GridDev G, const StageArgs A) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int strip = blockIdx.x * kBlockWaves + wave;

    double speed = 0.0;
    int bad = 0;

    if (strip < A.nstrips) {
        const int ng = G.ng, nx = G.nx, ny = G.ny, s = G.stride;
        const double half_g = G.half_g, g = G.g, inv_dx = G.inv_dx, inv_dy = G.inv_dy;

        const int j = ng - 2 + strip * kStripCols + lane;
        const bool owner = lane >= 2 && lane < kWave - 2 && j < ng + ny;
        bool gcol;
        const int jc = mirror<BC>(j < ng + ny + 2 ? j : ng + ny + 1, ng, ng + ny, gcol);

        const int r0 = ng + blockIdx.y * A.tile_rows;
        const int r1 = (r0 + A.tile_rows < ng + nx) ? r0 + A.tile_rows : ng + nx;
        const int rmax = ng + nx + 1;

        double dt = 0.0;
        if (STAGE != 0) dt = *A.dt;

        auto load = [&](int r) -> Cell {
            bool gr;
            const int rs = mirror<BC>(r, ng, ng + nx, gr);
            return load_cell<BC>(A, rs * s, gr, jc, gcol);
        };

        // ---- prologue: rows r0-2 .. r0+2, x face r0-1/2
        const Cell P2 = load(r0 - 2);
        const Cell P1 = load(r0 - 1);
        Cell Ac = load(r0);
        Cell Bc = load(r0 + 1);
        Cell Nc = load(r0 + 2 <= rmax ? r0 + 2 : rmax);

        double u0h = 0.0, u0u = 0.0, u0v = 0.0;
        if (STAGE >= 2) {
            const int k = r0 * s + jc;
            u0h = A.u0_h[k];
            u0u = A.u0_hu[k];
            u0v = A.u0_hv[k];
        }

        // x differences d_{r0-3/2}, d_{r0-1/2}, d_{r0+1/2}
        const double e0 = P1.eta - P2.eta, u0d = P1.hu - P2.hu, v0d = P1.hv - P2.hv;
        const double e1 = Ac.eta - P1.eta, u1d = Ac.hu - P1.hu, v1d = Ac.hv - P1.hv;
        double e2 = Bc.eta - Ac.eta, u2d = Bc.hu - Ac.hu, v2d = Bc.hv - Ac.hv;

        // minus-side reconstruction of cell r0-1 at face r0-1/2
        const double sP_e = 0.5 * minmod(e0, e1);
        const double sP_u = 0.5 * minmod(u0d, u1d);
        const double sP_v = 0.5 * minmod(v0d, v1d);
        // slopes of cell r0: plus side at r0-1/2, minus side at r0+1/2
        double sA_e = 0.5 * minmod(e1, e2);
        double sA_u = 0.5 * minmod(u1d, u2d);
        double sA_v = 0.5 * minmod(v1d, v2d);

        FaceFlux fx = face_flux(half_g, g, P1.eta + sP_e, P1.hu + sP_u, P1.hv + sP_v, P1.b,
                                Ac.eta - sA_e, Ac.hu - sA_u, Ac.hv - sA_v, Ac.b);
        // plus-side flux of face r-1/2 as seen by cell r
        double fxp_h = fx.fh, fxp_hu = fx.fn_p, fxp_hv = fx.ft;

        for (int r = r0; r < r1; ++r) {
            // prefetch row r+3 and U^n at row r+1
            const int rn = (r + 3 <= rmax) ? r + 3 : rmax;
            const Cell Nn = load(rn);
            double n0h = 0.0, n0u = 0.0, n0v = 0.0;
            if (STAGE >= 2) {
                const int k = ((r + 1 < r1) ? r + 1 : r) * s + jc;
                n0h = A.u0_h[k];
                n0u = A.u0_hu[k];
                n0v = A.u0_hv[k];
            }

            // ---- x: slope of cell r+1, face r+1/2
            const double e3 = Nc.eta - Bc.eta, u3d = Nc.hu - Bc.hu, v3d = Nc.hv - Bc.hv;
            const double sB_e = 0.5 * minmod(e2, e3);
            const double sB_u = 0.5 * minmod(u2d, u3d);
            const double sB_v = 0.5 * minmod(v2d, v3d);

            fx = face_flux(half_g, g, Ac.eta + sA_e, Ac.hu + sA_u, Ac.hv + sA_v, Ac.b,
                           Bc.eta - sB_e, Bc.hu - sB_u, Bc.hv - sB_v, Bc.b);

            // ---- y: neighbours along the lanes, face j+1/2 in lane j
            const double dy_e = from_next(Ac.eta) - Ac.eta;
            const double dy_u = from_next(Ac.hu) - Ac.hu;
            const double dy_v = from_next(Ac.hv) - Ac.hv;
            const double sy_e = 0.5 * minmod(from_prev(dy_e), dy_e);
            const double sy_u = 0.5 * minmod(from_prev(dy_u), dy_u);
            const double sy_v = 0.5 * minmod(from_prev(dy_v), dy_v);
            const double pe = from_next(Ac.eta - sy_e);
            const double pu = from_next(Ac.hu - sy_u);
            const double pv = from_next(Ac.hv - sy_v);
            const double pb = from_next(Ac.b);
            // y faces: normal momentum is hv, tangential hu
            const FaceFlux fy = face_flux(half_g, g, Ac.eta + sy_e, Ac.hv + sy_v, Ac.hu + sy_u,
                                          Ac.b, pe, pv, pu, pb);
            const double fyp_h = from_prev(fy.fh);
            const double fyp_hu = from_prev(fy.ft);
            const double fyp_hv = from_prev(fy.fn_p);

            // ---- divergence (omp_kernels.hpp:281-283)
            const double rh = divergence(inv_dx, inv_dy, fx.fh, fxp_h, fy.fh, fyp_h);
            const double rhu = divergence(inv_dx, inv_dy, fx.fn_m, fxp_hu, fy.ft, fyp_hu);
            const double rhv = divergence(inv_dx, inv_dy, fx.ft, fxp_hv, fy.fn_m, fyp_hv);

            if (owner) {
                const int k = r * s + j;
                if (STAGE == 0) {
                    A.out_h[k] = rh;
                    A.out_hu[k] = rhu;
                    A.out_hv[k] = rhv;
                } else {
                    if (STAGE == 1) {
                        u0h = Ac.h;
                        u0u = Ac.hu;
                        u0v = Ac.hv;
                    }
                    double hN = rk_combine<STAGE>(u0h, Ac.h, dt, rh);
                    double huN = rk_combine<STAGE>(u0u, Ac.hu, dt, rhu);
                    double hvN = rk_combine<STAGE>(u0v, Ac.hv, dt, rhv);
                    // enforce_positivity (omp_kernels.hpp:317-325)
                    if (hN < kFloor) {
                        hN = kFloor;
                        huN = 0.0;
                        hvN = 0.0;
                    }
                    if (!isfinite(hN) || !isfinite(huN) || !isfinite(hvN)) bad = 1;
                    if (SPEED) speed = cell_speed(g, hN, huN, hvN, speed);
                    A.out_h[k] = hN;
                    A.out_hu[k] = huN;
                    A.out_hv[k] = hvN;
                }
            }

            // ---- shift the x window
            fxp_h = fx.fh;
            fxp_hu = fx.fn_p;
            fxp_hv = fx.ft;
            Ac = Bc;
            Bc = Nc;
            Nc = Nn;
            sA_e = sB_e;
            sA_u = sB_u;
            sA_v = sB_v;
            e2 = e3;
            u2d = u3d;
            v2d = v3d;
            u0h = n0h;
            u0u = n0u;
            u0v = n0v;
        }
    }

    if (STAGE != 0) {
        if (__any(bad)) {
            if (lane == 0) atomicOr(A.flag, 1);
        }
    }
    if (SPEED) {
        __shared__ double red[kBlockWaves];
        for (int off = kWave / 2; off > 0; off >>= 1) speed = smax(speed, __shfl_xor(speed, off));
        if (lane == 0) red[wave] = speed;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = red[0];
            for (int w = 1; w < kBlockWaves; ++w) m = smax(m, red[w]);
            A.partial[blockIdx.y * gridDim.x + blockIdx.x] = m;
        }
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
                                                          long long hist_cap) {
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
        const double dt = dt_from_speed(r, cfl, dx, dy);
        scal[0] = dt;
        scal[1] = r;
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

__global__ void fill_double_kernel(double *dst, double v, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = v;
}

template <int STAGE, int BC, bool SPEED>
hipError_t launch_one(const GridDev &G, const StageArgs &A, hipStream_t stream, dim3 grid) {
    hipLaunchKernelGGL((stage_kernel<STAGE, BC, SPEED>), grid, dim3(kWave * kBlockWaves), 0,
                       stream, G, A);
    return hipGetLastError();
}

template <int BC>
hipError_t launch_bc(int stage, bool speed, const GridDev &G, const StageArgs &A,
                     hipStream_t stream, dim3 grid) {
    switch (stage) {
        case 0: return launch_one<0, BC, false>(G, A, stream, grid);
        case 1: return launch_one<1, BC, false>(G, A, stream, grid);
        case 2: return launch_one<2, BC, false>(G, A, stream, grid);
        case 3:
            return speed ? launch_one<3, BC, true>(G, A, stream, grid)
                         : launch_one<3, BC, false>(G, A, stream, grid);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

dim3 stage_grid(const GridDev &G, int tile_rows) {
    const int nstrips = strips_for(G.ny);
    return dim3((nstrips + kBlockWaves - 1) / kBlockWaves, (G.nx + tile_rows - 1) / tile_rows, 1);
}

hipError_t launch_stage(int stage, int bc, bool speed, const GridDev &G, const StageArgs &A,
                        hipStream_t stream, dim3 *grid_out) {
    if (A.tile_rows < 1 || A.nstrips != strips_for(G.ny)) return hipErrorInvalidValue;
    const dim3 grid = stage_grid(G, A.tile_rows);
    if (grid_out) *grid_out = grid;
    return bc == 0 ? launch_bc<0>(stage, speed, G, A, stream, grid)
                   : launch_bc<1>(stage, speed, G, A, stream, grid);
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
                              hipStream_t stream) {
    hipLaunchKernelGGL(finalize_dt_kernel, dim3(1), dim3(256), 0, stream, partial, n, cfl, dx, dy,
                       scal, hist, counter, hist_cap);
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

hipError_t launch_fill_double(double *dst, double v, long n, hipStream_t stream) {
    hipLaunchKernelGGL(fill_double_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       dst, v, n);
    return hipGetLastError();
}

}  // namespace swe
