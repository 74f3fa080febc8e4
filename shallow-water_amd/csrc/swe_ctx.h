// swe_ctx.h -- the single-grid context behind the C ABI (shared by the step and the
// adjoint entry points).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/swe_hip.h"
#include "swe_internal.h"

struct swe_hip_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    swe::GridDev G{};
    int bc = 0;
    size_t n_total = 0;
    double *U[3] = {nullptr, nullptr, nullptr};
    double *U1[3] = {nullptr, nullptr, nullptr};
    double *U2[3] = {nullptr, nullptr, nullptr};
    // fused steps (one launch per step): the state alternates between U and Unext, swapped
    // after every launch so that U is always the current state; fused_rows = rows per tile,
    // 0 = off
    double *Unext[3] = {nullptr, nullptr, nullptr};
    int fused_rows = 0;
    double *B = nullptr;
    // the bottom is +0.0 in every cell, ghosts included (checked at every upload); with
    // flat_bottom_mode on (the default) the stage launches then take the kernels that do not
    // read it
    bool b_zero = false;
    int flat_bottom_mode = 1;
    int *bz = nullptr;  // device word of that check
    double *scal = nullptr;      // [0] dt, [1] max speed
    double *partial = nullptr;   // per-block maxima
    int partial_cap = 0;
    int *flag = nullptr;         // run status words (swe::kStatusWords)
    int64_t last_status_step = 0;  // of the last stepping call: steps it started
    int last_status_stage = 0;     // stage whose output was non-finite, 0 if none
    long long *counter = nullptr;
    double *hist = nullptr;
    long long hist_cap = 0;
    // interior rows per wave of the stage launches, per stage (stage 0, the right-hand side
    // alone, takes stage 1's)
    int stage_rows[3] = {64, 64, 64};
    int precision = SWE_PRECISION_EXACT;
    int scheme = SWE_SCHEME_OPTIMIZED;  // swe_scheme; recon / riemann for the serial family
    int recon = SWE_RECON_MUSCL;
    int riemann = SWE_RIEMANN_HLL;
    bool use_graph = true;
    // cached graph of one step: key = (mode, stage tile heights, stream, hist pointer)
    hipGraphExec_t graph = nullptr;
    int graph_mode = -1;
    int graph_rows[3] = {-1, -1, -1};
    int graph_flat = -1;
    int graph_precision = -1;
    int graph_scheme = -1;  // scheme * 16 + recon * 4 + riemann of the captured step
    hipStream_t graph_stream = nullptr;
    double graph_cfl = 0.0;
    std::string err;
};
