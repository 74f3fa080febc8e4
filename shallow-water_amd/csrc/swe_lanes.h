// swe_lanes.h -- cross-lane moves, wave votes, ghost-index mirroring and row-based
// addressing shared by the stage, modular-scheme and adjoint kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace swe {

// Lane l receives lane l+1's value / lane l-1's value, through the DPP wavefront shifts
// wave_shl:1 (0x130) and wave_shr:1 (0x138) of the GFX9 ISA: a VALU move, no LDS round
// trip.  bound_ctrl is set, so the lane without a source (63 resp. 0) receives 0: one
// v_mov_b32_dpp per half (a form that keeps the destination lane's own value would first
// copy the source into the destination, two more moves per double).  Callers must not let
// lane 63 (from_next_z) or lane 0 (from_prev_z) reach a result, a store or a wave-wide
// vote (see wave_all's skip).
template <int CTRL>
__device__ __forceinline__ double dpp_move_z(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_next_z(double x) { return dpp_move_z<0x130>(x); }
__device__ __forceinline__ double from_prev_z(double x) { return dpp_move_z<0x138>(x); }

// Lanes 62 and 63: the lane whose from_next_z value has no source, and the lane whose
// y face reads that lane's slope.  Both are halo columns of the stage kernels.
constexpr unsigned long long kLastLanes = 3ull << 62;

// True when p holds in every active lane outside `skip` (wave-uniform).  The vote is a
// ballot of the compare's lane mask against exec, all on the scalar unit; HIP's __all(int)
// first widens the mask to a 0/1 VGPR and compares it back, two VALU instructions a vote.
__device__ __forceinline__ bool wave_all(bool p, unsigned long long skip = 0ull) {
    const unsigned long long exec = __builtin_amdgcn_read_exec();
    return ((__builtin_amdgcn_ballot_w64(p) | skip) & exec) == exec;
}

// True when p holds in some active lane (wave-uniform), on the scalar unit.
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// wave_all(p && q, skip) with one ballot per compare, combined on the scalar unit (LLVM
// otherwise materializes the && of two compares as a 0/1 VGPR before the vote).
__device__ __forceinline__ bool wave_all2(bool p, bool q, unsigned long long skip = 0ull) {
    const unsigned long long exec = __builtin_amdgcn_read_exec();
    return (((__builtin_amdgcn_ballot_w64(p) & __builtin_amdgcn_ballot_w64(q)) | skip) & exec) ==
           exec;
}

// wave_any(p || q || r), one ballot per compare.
__device__ __forceinline__ bool wave_any3(bool p, bool q, bool r) {
    return (__builtin_amdgcn_ballot_w64(p) | __builtin_amdgcn_ballot_w64(q) |
            __builtin_amdgcn_ballot_w64(r)) != 0;
}

// Value of lane `src` (wave-uniform) in every lane.
__device__ __forceinline__ double read_lane(double x, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
    return __hiloint2double(hi, lo);
}

// Source index of row/column r for the interior range lo..hi-1: r itself inside, the
// mirrored cell 2*lo-1-r / 2*hi-1-r for reflecting walls (BC 0), the nearest interior cell
// for transmissive ones (BC 1).  r may lie up to 2 cells outside the interior.
template <int BC>
__device__ __forceinline__ int mirror(int r, int lo, int hi, bool &ghost) {
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

// The row offset is wave-uniform and the column is fixed per lane, so every access is a
// scalar row base plus a constant 32-bit per-lane byte offset: the global_load/store
// "saddr + voffset" form, with no per-row 64-bit address arithmetic on the vector ALUs.
// row_off = instance offset (0 for a single grid) + row * stride, wave-uniform.  The row
// base passes through an empty asm statement that pins it to scalar registers: without
// it LLVM reassociates (base + row) + lane into a hoisted per-lane 64-bit pointer
// (two VGPRs per array) plus a 64-bit vector add per access.
__device__ __forceinline__ unsigned long long uniform_addr(unsigned long long p) {
    asm("" : "+s"(p));
    return p;
}
// Pointers into global memory (address space 1): loads and stores through them are
// global_* instructions, which have the saddr form (flat_* ones do not).
using GConstRow = const __attribute__((address_space(1))) char *;
using GRow = __attribute__((address_space(1))) char *;
__device__ __forceinline__ GConstRow row_ptr(const double *base, long long row_off) {
    return (GConstRow)uniform_addr(reinterpret_cast<unsigned long long>(base + row_off));
}
__device__ __forceinline__ GRow row_ptr(double *base, long long row_off) {
    return (GRow)uniform_addr(reinterpret_cast<unsigned long long>(base + row_off));
}
// Re-define a per-lane byte offset in place at row r, with no instruction.  The offset
// stays loop-variant, so LLVM does not hoist its zero extension out of the row loop: a
// hoisted 64-bit extension reaches the loop's loads and stores as a plain register pair,
// which rules out the saddr + voffset form (one 64-bit vector add per access instead).
// Call it in the basic block of the accesses that use the offset.
__device__ __forceinline__ void pin_lane_offset(unsigned &off, int r) {
    asm("" : "+v"(off) : "s"(r));
}
__device__ __forceinline__ double ld(GConstRow row, unsigned byte_off) {
    return *(const __attribute__((address_space(1))) double *)(row + byte_off);
}
__device__ __forceinline__ void st(GRow row, unsigned byte_off, double v) {
    *(__attribute__((address_space(1))) double *)(row + byte_off) = v;
}

}  // namespace swe
