// swe_batch_api.hip -- C ABI of the batched ensemble (declared in include/swe_hip.h).
//
// An ensemble holds `count` independent grids of uneven size on one device.  Each field
// (h, hu, hv, their two stage copies, B) is one slab; instance i occupies a 256-byte
// aligned range in the reference layout of its own grid.  A step is three launches of the
// batched stage kernel over a tile work-list (one tile = one wave's 60-column strip of
// tile_rows rows of one instance) -- every instance advances in the same launches, each
// with its own dt -- plus, in adaptive mode, one 1-block launch that turns the per-instance
// max wave speeds into the next dts.  One step is captured in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/swe_hip.h"
#include "swe_internal.h"

using swe::BatchArgs;
using swe::BatchInst;
using swe::BatchTile;

struct swe_hip_batch {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int count = 0;
    int bc = 0;
    double g = 9.80665;
    int precision = SWE_PRECISION_EXACT;
    int num_cus = 256;
    std::vector<swe_grid_desc> desc;
    std::vector<BatchInst> inst_h;
    BatchInst *inst_d = nullptr;
    long long total = 0;  // elements per slab
    double *U[3] = {nullptr, nullptr, nullptr};
    double *U1[3] = {nullptr, nullptr, nullptr};
    double *U2[3] = {nullptr, nullptr, nullptr};
    double *B = nullptr;
    double *dt_d = nullptr;
    unsigned long long *max_bits = nullptr;
    int *flags = nullptr;
    long long *counter = nullptr;
    double *hist = nullptr;
    long long hist_rows = 0;
    std::vector<BatchTile> tiles_h;
    BatchTile *tiles_d = nullptr;
    int ntiles = 0;
    int tile_rows = 32;
    int waves_per_cu = 0;  // 0: one tile per wave
    bool use_graph = true;
    hipGraphExec_t graph = nullptr;
    int graph_mode = -1;
    double graph_cfl = 0.0;
    std::string err;
};

namespace {

int bfail(swe_hip_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    return code;
}

#define SWE_BCHECK(b, call)                                                                    \
    do {                                                                                       \
        const hipError_t e__ = (call);                                                         \
        if (e__ != hipSuccess)                                                                 \
            return bfail(b, SWE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e__)); \
    } while (0)

void drop_graph(swe_hip_batch *b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_mode = -1;
}

int blocks_for(const swe_hip_batch *b) {
    const int bw = swe::kBatchBlockWaves;
    const int all = (b->ntiles + bw - 1) / bw;
    if (b->waves_per_cu <= 0) return all;
    const int want = (b->waves_per_cu * b->num_cus + bw - 1) / bw;
    return std::max(1, std::min(all, want));
}

// Work-list for the current tile_rows, instance-major so neighbouring waves share halo rows.
// The work list: full-height tiles first (instance by instance, row segments in order),
// then each instance's shorter last row segment, longest first, so the short tiles fill
// the end of the launch.  Each tile computes the same values in any order.
int build_tiles(swe_hip_batch *b) {
    b->tiles_h.clear();
    std::vector<std::pair<int, BatchTile>> tails;  // (rows, tile)
    for (int i = 0; i < b->count; ++i) {
        const int strips = swe::strips_for(b->inst_h[i].ny);
        const int nx = b->inst_h[i].nx;
        for (int r0 = 2; r0 < 2 + nx; r0 += b->tile_rows) {
            const int rows = std::min(b->tile_rows, 2 + nx - r0);
            for (int s = 0; s < strips; ++s) {
                if (rows == b->tile_rows)
                    b->tiles_h.push_back(BatchTile{i, s, r0, 0});
                else
                    tails.push_back({rows, BatchTile{i, s, r0, 0}});
            }
        }
    }
    std::stable_sort(tails.begin(), tails.end(),
                     [](const auto &x, const auto &y) { return x.first > y.first; });
    for (const auto &t : tails) b->tiles_h.push_back(t.second);
    drop_graph(b);
    if (b->tiles_d) (void)hipFree(b->tiles_d);
    b->tiles_d = nullptr;
    b->ntiles = (int)b->tiles_h.size();
    SWE_BCHECK(b, hipMalloc(&b->tiles_d, sizeof(BatchTile) * b->tiles_h.size()));
    SWE_BCHECK(b, hipMemcpyAsync(b->tiles_d, b->tiles_h.data(), sizeof(BatchTile) * b->tiles_h.size(),
                                 hipMemcpyHostToDevice, b->stream));
    SWE_BCHECK(b, hipStreamSynchronize(b->stream));
    return SWE_OK;
}

BatchArgs batch_args(swe_hip_batch *b, int stage) {
    BatchArgs a{};
    double *const *in = stage <= 1 ? b->U : (stage == 2 ? b->U1 : b->U2);
    double *const *out = stage == 1 ? b->U1 : (stage == 2 ? b->U2 : b->U);
    a.base.in_h = in[0];
    a.base.in_hu = in[1];
    a.base.in_hv = in[2];
    a.base.u0_h = b->U[0];
    a.base.u0_hu = b->U[1];
    a.base.u0_hv = b->U[2];
    a.base.B = b->B;
    a.base.out_h = out[0];
    a.base.out_hu = out[1];
    a.base.out_hv = out[2];
    a.base.tile_rows = b->tile_rows;
    a.inst = b->inst_d;
    a.tiles = b->tiles_d;
    a.ntiles = b->ntiles;
    a.dt = b->dt_d;
    a.max_bits = b->max_bits;
    a.flags = b->flags;
    a.g = b->g;
    a.half_g = 0.5 * b->g;
    a.split_ok = (b->g >= 0x1p-10 && b->g <= 0x1p10) ? 1 : 0;
    return a;
}

int enqueue_batch_step(swe_hip_batch *b, int mode, double cfl) {
    const bool adaptive = mode == 1;
    const int blocks = blocks_for(b);
    for (int stage = 1; stage <= 3; ++stage) {
        const BatchArgs a = batch_args(b, stage);
        SWE_BCHECK(b, swe::launch_batch_stage(stage, b->bc, adaptive && stage == 3,
                                              b->precision == SWE_PRECISION_FAST, a, blocks,
                                              b->stream));
    }
    if (adaptive)
        SWE_BCHECK(b, swe::launch_batch_finalize(b->inst_d, b->count, b->max_bits, cfl, b->dt_d,
                                                 b->hist, b->counter, b->hist_rows, b->flags,
                                                 b->stream));
    return SWE_OK;
}

int run_batch_steps(swe_hip_batch *b, int mode, double cfl, int64_t nsteps) {
    if (!b->use_graph || nsteps < 2) {
        for (int64_t s = 0; s < nsteps; ++s) {
            const int rc = enqueue_batch_step(b, mode, cfl);
            if (rc) return rc;
        }
        return SWE_OK;
    }
    if (!b->graph || b->graph_mode != mode || (mode == 1 && b->graph_cfl != cfl)) {
        drop_graph(b);
        SWE_BCHECK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_batch_step(b, mode, cfl);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        SWE_BCHECK(b, e);
        const hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SWE_BCHECK(b, ei);
        b->graph_mode = mode;
        b->graph_cfl = cfl;
    }
    for (int64_t s = 0; s < nsteps; ++s) SWE_BCHECK(b, hipGraphLaunch(b->graph, b->stream));
    return SWE_OK;
}

// Per instance: 0, or the stage whose output was non-finite (the instance stopped there),
// from the run status words (swe::kStatusWords per instance).
int read_outcomes(swe_hip_batch *b, std::vector<int> &stage, std::vector<int> &step) {
    std::vector<int> w((size_t)b->count * swe::kStatusWords);
    SWE_BCHECK(b, hipMemcpyAsync(w.data(), b->flags, sizeof(int) * w.size(),
                                 hipMemcpyDeviceToHost, b->stream));
    SWE_BCHECK(b, hipStreamSynchronize(b->stream));
    stage.assign((size_t)b->count, 0);
    step.assign((size_t)b->count, 0);
    for (size_t i = 0; i < (size_t)b->count; ++i) {
        const int *st = &w[i * swe::kStatusWords];
        step[i] = st[0];
        for (int s = 1; s <= 3; ++s)
            if (st[s] & swe::kStatusBad) stage[i] = s;
    }
    return SWE_OK;
}

int check_flags(swe_hip_batch *b) {
    std::vector<int> stage, step;
    const int rc = read_outcomes(b, stage, step);
    if (rc) return rc;
    for (int i = 0; i < b->count; ++i)
        if (stage[(size_t)i])
            return bfail(b, SWE_ERR_NONFINITE,
                         "non-finite state detected in ensemble instance " + std::to_string(i) +
                             " at stage " + std::to_string(stage[(size_t)i]) + " of step " +
                             std::to_string(step[(size_t)i]) + " of this call");
    return SWE_OK;
}

swe::GridDev inst_griddev(const swe_hip_batch *b, int i) {
    const BatchInst &I = b->inst_h[(size_t)i];
    return swe::GridDev{I.nx, I.ny, 2, I.stride, I.dx, I.dy, I.inv_dx, I.inv_dy, b->g,
                        0.5 * b->g, 0};
}

int ensure_hist(swe_hip_batch *b, long long rows) {
    if (b->hist_rows >= rows) return SWE_OK;
    drop_graph(b);
    if (b->hist) (void)hipFree(b->hist);
    b->hist = nullptr;
    b->hist_rows = 0;
    SWE_BCHECK(b, hipMalloc(&b->hist, sizeof(double) * (size_t)rows * (size_t)b->count));
    b->hist_rows = rows;
    return SWE_OK;
}

// Per-instance max wave speed of U -> dt (history row when counting).
int enqueue_initial_dt(swe_hip_batch *b, double cfl, bool record) {
    // record: the start of an adaptive step call, whose status words were just cleared
    SWE_BCHECK(b, swe::launch_fill_int(reinterpret_cast<int *>(b->max_bits), 0, 2 * b->count,
                                       b->stream));
    const BatchArgs a = batch_args(b, 1);
    SWE_BCHECK(b, swe::launch_batch_speed(a, blocks_for(b), b->stream));
    SWE_BCHECK(b, swe::launch_batch_finalize(b->inst_d, b->count, b->max_bits, cfl, b->dt_d,
                                             record ? b->hist : nullptr,
                                             record ? b->counter : nullptr, b->hist_rows,
                                             record ? b->flags : nullptr, b->stream));
    return SWE_OK;
}

}  // namespace

extern "C" {

int swe_hip_batch_create(const swe_grid_desc *descs, int32_t count, int device,
                         swe_hip_batch **out) {
    if (!descs || !out || count < 1) return SWE_ERR_INVALID;
    *out = nullptr;
    for (int i = 0; i < count; ++i) {
        const swe_grid_desc &d = descs[i];
        if (d.ng != 2 || d.nx < 2 || d.ny < 2 || !(d.dx > 0.0) || !(d.dy > 0.0) ||
            d.g != descs[0].g || !(d.g > 0.0) || d.bc_kind != descs[0].bc_kind ||
            (d.bc_kind != SWE_BC_REFLECTING && d.bc_kind != SWE_BC_TRANSMISSIVE))
            return SWE_ERR_INVALID;
        if ((long long)(d.nx + 4) * (d.ny + 4) >= (1LL << 31)) return SWE_ERR_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SWE_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return SWE_ERR_INVALID;

    auto *b = new swe_hip_batch();
    auto bad = [&](int code) {
        swe_hip_batch_destroy(b);
        return code;
    };
    b->device = device;
    b->count = count;
    b->bc = descs[0].bc_kind;
    b->g = descs[0].g;
    b->desc.assign(descs, descs + count);
    if (hipSetDevice(device) != hipSuccess) return bad(SWE_ERR_HIP);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        b->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking) != hipSuccess)
        return bad(SWE_ERR_HIP);
    b->stream = b->own_stream;

    long long off = 0;
    for (int i = 0; i < count; ++i) {
        const swe_grid_desc &d = descs[i];
        BatchInst I{};
        I.nx = d.nx;
        I.ny = d.ny;
        I.stride = d.ny + 4;
        I.offset = off;
        I.dx = d.dx;
        I.dy = d.dy;
        I.inv_dx = 1.0 / d.dx;
        I.inv_dy = 1.0 / d.dy;
        b->inst_h.push_back(I);
        const long long n = (long long)(d.nx + 4) * (d.ny + 4);
        off += (n + 31) / 32 * 32;  // 256-byte aligned instance starts
    }
    b->total = off;
    const size_t bytes = sizeof(double) * (size_t)b->total;
    double **arrays[] = {&b->U[0], &b->U[1], &b->U[2], &b->U1[0], &b->U1[1],
                         &b->U1[2], &b->U2[0], &b->U2[1], &b->U2[2], &b->B};
    for (double **a : arrays) {
        if (hipMalloc(a, bytes) != hipSuccess) return bad(SWE_ERR_OOM);
        if (swe::launch_fill_double(*a, 0.0, b->total, b->stream) != hipSuccess)
            return bad(SWE_ERR_HIP);
    }
    if (hipMalloc(&b->inst_d, sizeof(BatchInst) * (size_t)count) != hipSuccess ||
        hipMalloc(&b->dt_d, sizeof(double) * (size_t)count) != hipSuccess ||
        hipMalloc(&b->max_bits, sizeof(unsigned long long) * (size_t)count) != hipSuccess ||
        hipMalloc(&b->flags, sizeof(int) * swe::kStatusWords * (size_t)count) != hipSuccess ||
        hipMalloc(&b->counter, sizeof(long long)) != hipSuccess)
        return bad(SWE_ERR_OOM);
    if (hipMemcpyAsync(b->inst_d, b->inst_h.data(), sizeof(BatchInst) * (size_t)count,
                       hipMemcpyHostToDevice, b->stream) != hipSuccess)
        return bad(SWE_ERR_HIP);
    if (swe::launch_fill_double(b->dt_d, 0.0, count, b->stream) != hipSuccess ||
        swe::launch_fill_int(b->flags, 0, count * swe::kStatusWords, b->stream) != hipSuccess ||
        swe::launch_fill_int(reinterpret_cast<int *>(b->max_bits), 0, 2 * count, b->stream) !=
            hipSuccess ||
        swe::launch_fill_int(reinterpret_cast<int *>(b->counter), 0, 2, b->stream) != hipSuccess)
        return bad(SWE_ERR_HIP);
    if (build_tiles(b) != SWE_OK) return bad(SWE_ERR_HIP);
    if (hipStreamSynchronize(b->stream) != hipSuccess) return bad(SWE_ERR_HIP);
    *out = b;
    return SWE_OK;
}

void swe_hip_batch_destroy(swe_hip_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (int i = 0; i < 3; ++i) {
        (void)hipFree(b->U[i]);
        (void)hipFree(b->U1[i]);
        (void)hipFree(b->U2[i]);
    }
    (void)hipFree(b->B);
    (void)hipFree(b->inst_d);
    (void)hipFree(b->dt_d);
    (void)hipFree(b->max_bits);
    (void)hipFree(b->flags);
    (void)hipFree(b->counter);
    (void)hipFree(b->hist);
    (void)hipFree(b->tiles_d);
    if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
    delete b;
}

const char *swe_hip_batch_last_error(const swe_hip_batch *b) { return b ? b->err.c_str() : ""; }

int swe_hip_batch_count(const swe_hip_batch *b) { return b ? b->count : -1; }

int swe_hip_batch_upload(swe_hip_batch *b, int32_t i, const double *h, const double *hu,
                         const double *hv, const double *bathy) {
    if (!b || i < 0 || i >= b->count || !h || !hu || !hv || !bathy) return SWE_ERR_INVALID;
    SWE_BCHECK(b, hipSetDevice(b->device));
    const BatchInst &I = b->inst_h[(size_t)i];
    const size_t bytes = sizeof(double) * (size_t)(I.nx + 4) * (size_t)(I.ny + 4);
    SWE_BCHECK(b, hipMemcpyAsync(b->U[0] + I.offset, h, bytes, hipMemcpyHostToDevice, b->stream));
    SWE_BCHECK(b, hipMemcpyAsync(b->U[1] + I.offset, hu, bytes, hipMemcpyHostToDevice, b->stream));
    SWE_BCHECK(b, hipMemcpyAsync(b->U[2] + I.offset, hv, bytes, hipMemcpyHostToDevice, b->stream));
    SWE_BCHECK(b, hipMemcpyAsync(b->B + I.offset, bathy, bytes, hipMemcpyHostToDevice, b->stream));
    const swe::GridDev G = inst_griddev(b, i);
    SWE_BCHECK(b, swe::launch_ghost_fill(b->bc, G, b->B + I.offset, nullptr, nullptr, b->stream));
    SWE_BCHECK(b, swe::launch_ghost_fill(b->bc, G, b->U[0] + I.offset, b->U[1] + I.offset,
                                         b->U[2] + I.offset, b->stream));
    SWE_BCHECK(b, hipStreamSynchronize(b->stream));
    return SWE_OK;
}

int swe_hip_batch_download(swe_hip_batch *b, int32_t i, double *h, double *hu, double *hv) {
    if (!b || i < 0 || i >= b->count || !h || !hu || !hv) return SWE_ERR_INVALID;
    SWE_BCHECK(b, hipSetDevice(b->device));
    const BatchInst &I = b->inst_h[(size_t)i];
    const size_t bytes = sizeof(double) * (size_t)(I.nx + 4) * (size_t)(I.ny + 4);
    SWE_BCHECK(b, swe::launch_ghost_fill(b->bc, inst_griddev(b, i), b->U[0] + I.offset,
                                         b->U[1] + I.offset, b->U[2] + I.offset, b->stream));
    SWE_BCHECK(b, hipMemcpyAsync(h, b->U[0] + I.offset, bytes, hipMemcpyDeviceToHost, b->stream));
    SWE_BCHECK(b, hipMemcpyAsync(hu, b->U[1] + I.offset, bytes, hipMemcpyDeviceToHost, b->stream));
    SWE_BCHECK(b, hipMemcpyAsync(hv, b->U[2] + I.offset, bytes, hipMemcpyDeviceToHost, b->stream));
    SWE_BCHECK(b, hipStreamSynchronize(b->stream));
    return SWE_OK;
}

int swe_hip_batch_set_dt(swe_hip_batch *b, const double *dt) {
    if (!b || !dt) return SWE_ERR_INVALID;
    for (int i = 0; i < b->count; ++i)
        if (!(dt[i] > 0.0)) return SWE_ERR_INVALID;
    SWE_BCHECK(b, hipSetDevice(b->device));
    SWE_BCHECK(b, hipMemcpyAsync(b->dt_d, dt, sizeof(double) * (size_t)b->count,
                                 hipMemcpyHostToDevice, b->stream));
    SWE_BCHECK(b, hipStreamSynchronize(b->stream));
    return SWE_OK;
}

int swe_hip_batch_stable_dt(swe_hip_batch *b, double cfl, double *dt_out) {
    if (!b || !(cfl > 0.0)) return SWE_ERR_INVALID;
    SWE_BCHECK(b, hipSetDevice(b->device));
    int rc = enqueue_initial_dt(b, cfl, false);
    if (rc) return rc;
    if (dt_out)
        SWE_BCHECK(b, hipMemcpyAsync(dt_out, b->dt_d, sizeof(double) * (size_t)b->count,
                                     hipMemcpyDeviceToHost, b->stream));
    SWE_BCHECK(b, hipStreamSynchronize(b->stream));
    return SWE_OK;
}

int swe_hip_batch_step(swe_hip_batch *b, int64_t nsteps) {
    if (!b || nsteps < 0) return SWE_ERR_INVALID;
    SWE_BCHECK(b, hipSetDevice(b->device));
    SWE_BCHECK(b, swe::launch_fill_int(b->flags, 0, b->count * swe::kStatusWords, b->stream));
    const int rc = run_batch_steps(b, 0, 0.0, nsteps);
    if (rc) return rc;
    return check_flags(b);
}

int swe_hip_batch_step_adaptive(swe_hip_batch *b, double cfl, int64_t nsteps, double *dt_hist) {
    if (!b || nsteps < 0 || !(cfl > 0.0)) return SWE_ERR_INVALID;
    SWE_BCHECK(b, hipSetDevice(b->device));
    int rc = ensure_hist(b, nsteps + 1);
    if (rc) return rc;
    SWE_BCHECK(b, swe::launch_fill_int(b->flags, 0, b->count * swe::kStatusWords, b->stream));
    SWE_BCHECK(b, swe::launch_fill_int(reinterpret_cast<int *>(b->counter), 0, 2, b->stream));
    rc = enqueue_initial_dt(b, cfl, true);
    if (rc) return rc;
    rc = run_batch_steps(b, 1, cfl, nsteps);
    if (rc) return rc;
    if (dt_hist && nsteps > 0)
        SWE_BCHECK(b, hipMemcpyAsync(dt_hist, b->hist,
                                     sizeof(double) * (size_t)nsteps * (size_t)b->count,
                                     hipMemcpyDeviceToHost, b->stream));
    return check_flags(b);
}

int swe_hip_batch_nonfinite(swe_hip_batch *b, int32_t *flags) {
    if (!b || !flags) return SWE_ERR_INVALID;
    std::vector<int> stage, step;
    const int rc = read_outcomes(b, stage, step);
    if (rc) return rc;
    for (int i = 0; i < b->count; ++i) flags[i] = stage[(size_t)i];
    return SWE_OK;
}

int swe_hip_batch_set_launch(swe_hip_batch *b, int tile_rows, int waves_per_cu) {
    if (!b || tile_rows < 1 || waves_per_cu < 0) return SWE_ERR_INVALID;
    SWE_BCHECK(b, hipSetDevice(b->device));
    b->waves_per_cu = waves_per_cu;
    drop_graph(b);
    if (tile_rows != b->tile_rows || !b->tiles_d) {
        b->tile_rows = tile_rows;
        return build_tiles(b);
    }
    return SWE_OK;
}

int swe_hip_batch_get_launch(const swe_hip_batch *b, int *tile_rows, int *waves_per_cu) {
    if (!b) return SWE_ERR_INVALID;
    if (tile_rows) *tile_rows = b->tile_rows;
    if (waves_per_cu) *waves_per_cu = b->waves_per_cu;
    return SWE_OK;
}

int swe_hip_batch_set_precision(swe_hip_batch *b, int mode) {
    if (!b || (mode != SWE_PRECISION_EXACT && mode != SWE_PRECISION_FAST)) return SWE_ERR_INVALID;
    b->precision = mode;
    drop_graph(b);
    return SWE_OK;
}

int swe_hip_batch_set_graph(swe_hip_batch *b, int enabled) {
    if (!b) return SWE_ERR_INVALID;
    b->use_graph = enabled != 0;
    if (!b->use_graph) drop_graph(b);
    return SWE_OK;
}

int swe_hip_batch_autotune(swe_hip_batch *b, int *tile_rows_out, int *waves_per_cu_out) {
    if (!b) return SWE_ERR_INVALID;
    SWE_BCHECK(b, hipSetDevice(b->device));
    // Time stages 1 and 2 only: they read U and write the stage buffers, so the state the
    // caller uploaded is untouched.  dt comes from the current per-instance dt array.
    const int tiles_c[] = {8, 16, 32, 64, 96, 128};
    const int waves_c[] = {8, 12, 16, 24, 0};
    swe::EventSet ev(2);
    SWE_BCHECK(b, ev.status);
    hipEvent_t e0 = ev[0], e1 = ev[1];
    // every instance's stages run (no first-bad-stage stop left over from an earlier call)
    SWE_BCHECK(b, swe::launch_fill_int(b->flags, 0, b->count * swe::kStatusWords, b->stream));
    float best = 1e30f;
    int best_t = b->tile_rows, best_w = b->waves_per_cu;
    for (int t : tiles_c) {
        int rc = swe_hip_batch_set_launch(b, t, 0);
        if (rc) return rc;
        for (int w : waves_c) {
            b->waves_per_cu = w;
            const int blocks = blocks_for(b);
            float total = 0.f;
            for (int rep = 0; rep < 3; ++rep) {
                SWE_BCHECK(b, hipEventRecord(e0, b->stream));
                for (int stage = 1; stage <= 2; ++stage)
                    SWE_BCHECK(b, swe::launch_batch_stage(stage, b->bc, false,
                                                          b->precision == SWE_PRECISION_FAST,
                                                          batch_args(b, stage), blocks, b->stream));
                SWE_BCHECK(b, hipEventRecord(e1, b->stream));
                SWE_BCHECK(b, hipEventSynchronize(e1));
                float ms = 0.f;
                SWE_BCHECK(b, hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) total += ms;  // first repetition warms caches and code
            }
            if (total < best) {
                best = total;
                best_t = t;
                best_w = w;
            }
        }
    }
    int rc = swe_hip_batch_set_launch(b, best_t, best_w);
    if (rc) return rc;
    if (tile_rows_out) *tile_rows_out = best_t;
    if (waves_per_cu_out) *waves_per_cu_out = best_w;
    return SWE_OK;
}

int swe_hip_batch_sync(swe_hip_batch *b) {
    if (!b) return SWE_ERR_INVALID;
    SWE_BCHECK(b, hipStreamSynchronize(b->stream));
    return SWE_OK;
}

}  // extern "C"
