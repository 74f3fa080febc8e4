// swe_internal.h -- device-side data layout and kernel launchers shared by the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace swe {

// hipEvents created together and destroyed with the object, so that every return path of
// a timing routine releases them.  `status` is the first creation error (hipSuccess if all
// were created).
struct EventSet {
    std::vector<hipEvent_t> ev;
    hipError_t status = hipSuccess;
    explicit EventSet(size_t n) : ev(n, nullptr) {
        for (auto &e : ev) {
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) {
                e = nullptr;
                if (status == hipSuccess) status = r;
            }
        }
    }
    ~EventSet() {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    EventSet(const EventSet &) = delete;
    EventSet &operator=(const EventSet &) = delete;
    hipEvent_t operator[](size_t i) const { return ev[i]; }
};

// One grid in the reference layout (omp_kernels.hpp:22-35): row-major, k = i*stride + j,
// i = x index (slow), j = y index (contiguous), stride = ny + 2*ng.
struct GridDev {
    int nx, ny, ng, stride;
    double dx, dy, inv_dx, inv_dy, g, half_g;
    int split_ok;  // g in [2^-10, 2^10]: the split division/sqrt forms may be used
};

// Arguments of one fused RK stage launch.
//   in_*  : stage input U_s (read through the radius-2 cross stencil)
//   u0_*  : U^n (read pointwise; stages 2 and 3)
//   out_* : stage output (stage 3 may alias u0_*; in_* never aliases out_*)
struct StageArgs {
    const double *in_h, *in_hu, *in_hv;
    const double *u0_h, *u0_hu, *u0_hv;
    const double *B;
    double *out_h, *out_hu, *out_hv;
    const double *dt;   // device scalar
    double *partial;    // per-block max wave speed (SPEED launches), may be null otherwise
    int *flag;          // run status words (kStatusWords, below)
    int tile_rows;      // interior rows per wave
    int nstrips;        // 60-column strips across y
    int flat_b;         // host side: B is +0.0 in every cell, ghosts included (the launch
                        // then takes the kernels that do not read B)
};

constexpr int kWave = 64;
constexpr int kStripCols = kWave - 4;  // output columns per wave (2-cell halo each side)
constexpr int kBlockWaves = 4;
// Waves per workgroup of the stage and batch kernels.  Their waves share nothing (no LDS,
// no barrier in the march), so the workgroup size only decides how the dispatcher refills
// the CUs.  Overridable at build time for tuning sweeps.
#ifndef SWE_STAGE_BLOCK_WAVES
#define SWE_STAGE_BLOCK_WAVES 4
#endif
#ifndef SWE_BATCH_BLOCK_WAVES
#define SWE_BATCH_BLOCK_WAVES 4
#endif
constexpr int kStageBlockWaves = SWE_STAGE_BLOCK_WAVES;
constexpr int kBatchBlockWaves = SWE_BATCH_BLOCK_WAVES;

__host__ __device__ inline int strips_for(int ny) { return (ny + kStripCols - 1) / kStripCols; }

// Run status of one grid: kStatusWords ints in device memory, zeroed at the start of every
// stepping call.
//   [0]     steps started in this call (stage-1 launches that ran)
//   [1..3]  stage s: kStatusBad = its output held a non-finite value, kStatusSkipped = a
//           launch of stage s did nothing because an earlier stage had one.
// The reference throws at the first stage whose output is non-finite (omp_kernels.hpp:
// 322-330, from rk3_step, omp_solver.hpp:250-338), so its arrays stay as that stage left
// them: U^n untouched when stage 1 or 2 failed, the bad stage-3 output in U when stage 3
// did.  Here every stage launch first reads the words of the two OTHER stages and, when
// either is set, only marks its own word and returns: the launch that finds the bad value
// runs to its end (its own blocks never read its own word), the next stage sees it and
// skips, the one after sees both, and so on, so every later launch of the call leaves the
// arrays alone.
constexpr int kStatusWords = 4;
constexpr int kStatusBad = 1, kStatusSkipped = 2;

// The words of the two other stages (uniform): non-zero when an earlier stage of this run
// went non-finite.  Read before the launch's first store, so that the compiler can use
// scalar loads; the launch then marks its own word (stage_skip) and writes nothing.
template <int STAGE>
__device__ __forceinline__ int stage_stop_words(const int *status) {
    constexpr int a = STAGE % 3 + 1, b = (STAGE + 1) % 3 + 1;
    return status[a] | status[b];
}
template <int STAGE>
__device__ __forceinline__ void stage_skip(int *status) {
    if (threadIdx.x == 0) atomicOr(status + STAGE, kStatusSkipped);
}

// stage: 0 = right-hand side only (out = rhs), 1..3 = SSP-RK3 stage with floor.
// bc: 0 = reflecting walls (reference), 1 = transmissive (zero-gradient).
// speed: also write per-block max wave speed of the output (stage 3 only).
// fast: reciprocal-based velocities (few-ulp differences from the reference arithmetic).
hipError_t launch_stage(int stage, int bc, bool speed, bool fast, const GridDev &G,
                        const StageArgs &A, hipStream_t stream, dim3 *grid_out);
// Grid of a strip x row-segment launch with `block_waves` waves (strips) per workgroup.
dim3 stage_grid(const GridDev &G, int tile_rows, int block_waves = kStageBlockWaves);

// One launch per SSP-RK3 step (the three stages fused, temporal blocking): A1 = stage 1
// (in U^n, out U1), A2 = stage 2 (in U1, u0 U^n, out U2), A3 = stage 3 (in U2, u0 U^n, out
// U^{n+1}, a fourth array set: never U^n); A1 carries dt, the status words, tile_rows (>= 4)
// and nstrips = fused_strips_for(ny) (52-column strips); parity = the launch's index in the
// call mod 2 (status words, see fused_step_kernel).  ng must be 2.  speed: per-block max wave
// speed of U^{n+1} into A3.partial (fused_step_blocks entries).
int fused_strips_for(int ny);
int fused_step_blocks(const GridDev &G, int tile_rows);
hipError_t launch_fused_step(int bc, bool speed, bool fast, const GridDev &G,
                             const StageArgs &A1, const StageArgs &A2, const StageArgs &A3,
                             int parity, hipStream_t stream, dim3 *grid_out);

// Modular serial scheme family (swe_modular.hip): bathy = the hydrostatic flux assembly
// (HLL only, as the reference dispatches it), else the plain one; recon 0 MUSCL / 1
// piecewise constant; riemann 0 HLL / 1 Rusanov / 2 ROE.  Reflecting walls, fixed dt.
bool modular_supported(bool bathy, int recon, int riemann);
hipError_t launch_modular_stage(int stage, bool bathy, int recon, int riemann, const GridDev &G,
                                const StageArgs &A, hipStream_t stream, dim3 *grid_out);

// Max wave speed over the interior of (h, hu, hv) into partial[0..blocks), blocks <= 1024.
hipError_t launch_speed_reduce(const GridDev &G, const double *h, const double *hu,
                               const double *hv, double *partial, int blocks,
                               hipStream_t stream);

// Reduce partial[0..n) to the max speed, then dt = compute_stable_dt expression.
//   scal[0] = dt, scal[1] = max speed.  If hist != null and *counter < hist_cap,
//   hist[*counter] = dt; *counter += 1.  When status (may be null) records a non-finite
//   stage, scal is left alone and the history entry is a quiet NaN.
hipError_t launch_finalize_dt(const double *partial, int n, double cfl, double dx, double dy,
                              double *scal, double *hist, long long *counter, long long hist_cap,
                              const int *status, hipStream_t stream);

// Ghost fill of (h, hu, hv) in place (hu == null: scalar fill of h only), reproducing the
// two-pass order of omp_solver.hpp:340-432 including the corner cells.
hipError_t launch_ghost_fill(int bc, const GridDev &G, double *h, double *hu, double *hv,
                             hipStream_t stream);

// ---- batched ensemble ---------------------------------------------------------------
// Instances of uneven size share one slab per field; instance i's arrays start at element
// offset[i] and have its own reference layout (ng = 2, stride = ny + 4).
struct BatchInst {
    int nx, ny, stride, pad;
    long long offset;
    double dx, dy, inv_dx, inv_dy;
};

// One entry of the tile work-list: instance, 60-column strip, first interior row.
struct BatchTile {
    int inst, strip, r0, pad;
};

struct BatchArgs {
    StageArgs base;               // slab pointers (in_*, u0_*, B, out_*); tile_rows
    const BatchInst *inst;
    const BatchTile *tiles;
    int ntiles;
    const double *dt;             // per instance
    unsigned long long *max_bits; // per instance, bits of a non-negative double (SPEED)
    int *flags;                   // per instance: kStatusWords run status words
    double g, half_g;
    int split_ok;
};

// Stage launch over the work-list: `blocks` blocks of kBatchBlockWaves waves; wave w takes
// tiles w, w + W, w + 2W, ... (W = kBatchBlockWaves*blocks), i.e. persistent when W < ntiles.
hipError_t launch_batch_stage(int stage, int bc, bool speed, bool fast, const BatchArgs &B,
                              int blocks, hipStream_t stream);
// Max wave speed of each instance's current state (U = base.in_*) into max_bits.
hipError_t launch_batch_speed(const BatchArgs &B, int blocks, hipStream_t stream);
// Per instance: dt = compute_stable_dt expression of max_bits (then zeroed); history row
// hist[*counter * count + i] when hist != null and *counter < hist_rows; counter += 1.  An
// instance whose status words (flags, kStatusWords each) record a non-finite stage keeps
// its dt and gets a quiet NaN history entry.
hipError_t launch_batch_finalize(const BatchInst *inst, int count, unsigned long long *max_bits,
                                 double cfl, double *dt, double *hist, long long *counter,
                                 long long hist_rows, const int *flags, hipStream_t stream);

hipError_t launch_set_scalar(double *dst, double v, hipStream_t stream);
hipError_t launch_fill_int(int *dst, int v, int n, hipStream_t stream);
hipError_t launch_fill_double(double *dst, double v, long n, hipStream_t stream);
// *flag |= 1 when some a[0..n) is not +0.0 (a bit pattern other than all zeros); the caller
// clears *flag first.
hipError_t launch_nonzero_bits(const double *a, long n, int *flag, hipStream_t stream);

// ---- discrete adjoint ------------------------------------------------------------------
// Arguments of one adjoint stage launch (stage s of a step, reversed).
//   v_*      : the stage input V (U^n, U1 or U2), read through the stencil
//   lam_*    : dL/d(stage output), interior cells
//   mask_out : 1 where the stage output is the floor state (1e-8, 0, 0)
//   out_*    : dL/dV (stage 1: dL/dU^n, the e_* accumulator included)
//   e_*      : dL/dU^n accumulator (stage 3 writes, stage 2 adds, stage 1 reads)
//   lam_b    : dL/dB accumulator (every stage adds)
//   mask_in  : written: 1 where V is the floor state (mask_out of the previous stage's adjoint)
struct AdjArgs {
    const double *v_h, *v_hu, *v_hv;
    const double *B;
    const double *lam_h, *lam_hu, *lam_hv;
    const unsigned char *mask_out;
    double *out_h, *out_hu, *out_hv;
    double *e_h, *e_hu, *e_hv;
    double *lam_b;
    unsigned char *mask_in;
    const double *dt;  // device scalar: this step's dt (frozen)
    int tile_rows;     // rows per wave (fused kernel) / x faces per wave (x-face kernel)
    int y_rows;        // rows per wave of the y-face kernel (two-kernel path)
    int block_waves;   // waves per block of the fused kernel (2, 4, 8); 0: two kernels
    int block_rows;    // fused kernel: most rows of a block (set by its launcher)
    int nstrips;       // 58-column strips across y (adj_strips_for)
    double *xf;        // two-kernel path: 8 x (nx+1) x stride adjoint outputs of the x faces
    unsigned char *xsel;  // two-kernel path: (nx+2) x stride minmod selections
};

int adj_strips_for(int ny);
// LDS of one fused-kernel block; launches above kAdjMaxLds are rejected.
constexpr size_t kAdjMaxLds = 160 * 1024;
size_t adj_lds_bytes(int block_waves, int tile_rows);
size_t adj_lds_rows_bytes(int block_rows);
bool adj_launch_ok(int block_waves, int tile_rows);
// ng must be 2.
hipError_t launch_adjoint_stage(int stage, int bc, const GridDev &G, const AdjArgs &A,
                                hipStream_t stream);
// mask[k] = 1 where (h, hu, hv)[k] is the floor state, 0 elsewhere (interior cells only).
hipError_t launch_floor_mask(const GridDev &G, const double *h, const double *hu,
                             const double *hv, unsigned char *mask, hipStream_t stream);
// Zero the ghost frame of one array.
hipError_t launch_zero_ghosts(const GridDev &G, double *a, hipStream_t stream);

}  // namespace swe
