// swe_api.hip -- C ABI of libswe_hip.so (declared in include/swe_hip.h).
//
// A context owns one grid's device arrays on one device and one stream:
//   U = (h, hu, hv), U1, U2 (stage buffers), B, and a few scalars (dt, max speed),
// the per-block partial maxima of the CFL reduction and the non-finite flag.
// A step is three fused stage launches U -> U1 -> U2 -> U (stage 3 writes U in place: it
// reads U^n only at the cell it writes).  With graphs enabled a run of kGraphSteps steps is
// captured once and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/swe_hip.h"
#include "swe_ctx.h"
#include "swe_internal.h"

using swe::GridDev;
using swe::StageArgs;


namespace {

int fail(swe_hip_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

// Message of the calling thread's last failed swe_hip_create (no context survives it).
thread_local std::string g_create_err;

#define SWE_CHECK(ctx, call)                                                                 \
    do {                                                                                     \
        const hipError_t e__ = (call);                                                       \
        if (e__ != hipSuccess) {                                                             \
            return fail(ctx, SWE_ERR_HIP,                                                    \
                        std::string(#call) + ": " + hipGetErrorString(e__) + " (" __FILE__ ":" + \
                            std::to_string(__LINE__) + ")");                                 \
        }                                                                                    \
    } while (0)

int default_tile_rows(const GridDev &G) {
    // 32 rows per wave measured best on 8192^2 (vs 64/128/256: fewer waves idle at the end
    // of the launch); smaller grids halve it, down to 4, until a launch holds two rounds of
    // the 3072 resident waves (256 CUs x 4 SIMDs x 3), trading the 4-row prologue of each
    // segment.  One round was the bar before the prologue's rows went out together: two
    // took 2048^2 from 16 to 8 rows (-1.8% per step) and 2560^2 from 32 / 24 to 16 rows
    // (-3.2%), 3072^2 to 16 rows (-0.6%) and 1536^2 from 8 to 4 rows (-1.3%); 256^2 -
    // 1024^2 and 4096^2 - 8192^2 unchanged (interleaved runs, one box).  The optimum is not
    // monotone in the size (1536^2 runs 16-row tiles 3.6% faster still), so callers that
    // step one size for long can time a few heights (swe_hip_set_stage_tile_rows).
    const long strips = swe::strips_for(G.ny);
    int tr = 32;
    while (tr > 4 && strips * ((G.nx + tr - 1) / tr) < 2L * 3 * 256 * 4) tr /= 2;
    return tr;
}

// Default tile heights per stage.  Stages 2 and 3 read U^n as well and run faster on
// shorter tiles than stage 1 does: on the 8192^2 dam break, stage 1 at 32 rows and stages
// 2, 3 at 24 beat 32 rows everywhere by 0.6-0.7% (interleaved runs on one box; 24 rows
// everywhere slows stage 1 by 2%).  Grids that already need shorter tiles keep one height.
void default_stage_rows(const GridDev &G, int rows[3]) {
    const int t = default_tile_rows(G);
    rows[0] = t;
    rows[1] = rows[2] = t >= 32 ? t * 3 / 4 : t;
}

void drop_graph(swe_hip_ctx *c) {
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
    c->graph_mode = -1;
}

int ensure_partial(swe_hip_ctx *c, int need) {
    if (need <= c->partial_cap) return SWE_OK;
    drop_graph(c);  // a captured step bakes the partial-max pointer
    if (c->partial) (void)hipFree(c->partial);
    c->partial = nullptr;
    c->partial_cap = 0;
    SWE_CHECK(c, hipMalloc(&c->partial, sizeof(double) * (size_t)need));
    c->partial_cap = need;
    return SWE_OK;
}

// The stage launches take the kernels that do not read B (same values).
bool flat_bottom(const swe_hip_ctx *c) { return c->flat_bottom_mode != 0 && c->b_zero; }

StageArgs stage_args(swe_hip_ctx *c, int stage) {
    StageArgs a{};
    double *const *in = stage == 1 || stage == 0 ? c->U : (stage == 2 ? c->U1 : c->U2);
    // stage 0 (rhs only) writes into the U1 buffers, never over its own input
    double *const *out = (stage == 0 || stage == 1) ? c->U1 : (stage == 2 ? c->U2 : c->U);
    a.in_h = in[0];
    a.in_hu = in[1];
    a.in_hv = in[2];
    a.u0_h = c->U[0];
    a.u0_hu = c->U[1];
    a.u0_hv = c->U[2];
    a.B = c->B;
    a.out_h = out[0];
    a.out_hu = out[1];
    a.out_hv = out[2];
    a.dt = c->scal;
    a.partial = c->partial;
    a.flag = c->flag;
    a.tile_rows = c->stage_rows[stage <= 1 ? 0 : stage - 1];
    a.flat_b = flat_bottom(c) ? 1 : 0;
    a.nstrips = swe::strips_for(c->G.ny);
    return a;
}

int scheme_key(const swe_hip_ctx *c) { return c->scheme * 16 + c->recon * 4 + c->riemann; }

// One stage launch of the context's scheme family.
hipError_t launch_ctx_stage(swe_hip_ctx *c, int stage, bool speed, const StageArgs &a,
                            dim3 *grid) {
    if (c->scheme == SWE_SCHEME_OPTIMIZED)
        return swe::launch_stage(stage, c->bc, speed, c->precision == SWE_PRECISION_FAST, c->G,
                                 a, c->stream, grid);
    return swe::launch_modular_stage(stage, c->scheme == SWE_SCHEME_SERIAL, c->recon, c->riemann,
                                     c->G, a, c->stream, grid);
}

// Enqueue one step.  mode 0: fixed dt (scal[0] holds it); mode 1: adaptive, stage 3 reduces
// the new max speed and a finalize kernel writes the next dt into scal[0] (and the history).
int enqueue_step(swe_hip_ctx *c, int mode, double cfl) {
    const bool speed = mode == 1;
    for (int stage = 1; stage <= 3; ++stage) {
        StageArgs a = stage_args(c, stage);
        dim3 grid;
        SWE_CHECK(c, launch_ctx_stage(c, stage, speed && stage == 3, a, &grid));
        if (speed && stage == 3) {
            SWE_CHECK(c, swe::launch_finalize_dt(c->partial, (int)(grid.x * grid.y), cfl, c->G.dx,
                                                 c->G.dy, c->scal, c->hist, c->counter,
                                                 c->hist_cap, c->flag, c->stream));
        }
    }
    return SWE_OK;
}

bool fused_mode(const swe_hip_ctx *c) {
    return c->fused_rows > 0 && c->scheme == SWE_SCHEME_OPTIMIZED;
}

// Fused steps: one launch per step (swe::launch_fused_step), U -> Unext, then the host swaps
// the two pointer sets so that U is the state the next launch reads.  Launched directly
// (the swap would change the pointers a captured graph holds).
int run_fused_steps(swe_hip_ctx *c, int mode, double cfl, int64_t nsteps) {
    drop_graph(c);  // a captured three-launch step holds the U pointers
    const int strips = swe::fused_strips_for(c->G.ny);
    for (int64_t s = 0; s < nsteps; ++s) {
        StageArgs a1 = stage_args(c, 1), a2 = stage_args(c, 2), a3 = stage_args(c, 3);
        a1.tile_rows = c->fused_rows;
        a1.nstrips = strips;
        a3.out_h = c->Unext[0];
        a3.out_hu = c->Unext[1];
        a3.out_hv = c->Unext[2];
        dim3 grid;
        SWE_CHECK(c, swe::launch_fused_step(c->bc, mode == 1,
                                            c->precision == SWE_PRECISION_FAST, c->G, a1, a2,
                                            a3, (int)(s & 1), c->stream, &grid));
        if (mode == 1)
            SWE_CHECK(c, swe::launch_finalize_dt(c->partial, (int)grid.x, cfl, c->G.dx, c->G.dy,
                                                 c->scal, c->hist, c->counter, c->hist_cap,
                                                 c->flag, c->stream));
        for (int i = 0; i < 3; ++i) std::swap(c->U[i], c->Unext[i]);
    }
    return SWE_OK;
}

// Steps per captured graph.  A graph launch costs the GPU 4-5 us beyond its kernels: a
// one-step graph replayed per step ran a 1024^2 step (85 us of kernels) 5.6% and a 256^2
// step 11% slower than direct launches, while 16- and 64-step graphs ran them 0.3-0.6% and
// 3-4% faster than direct launches (8192^2: all the same; interleaved runs, one box).
// Steps beyond the last whole graph launch directly; any call of 2 steps or more captures
// the graph if it is missing, so a short warm-up call prepares the long ones.
constexpr int kGraphSteps = 32;

int run_steps(swe_hip_ctx *c, int mode, double cfl, int64_t nsteps) {
    if (fused_mode(c)) return run_fused_steps(c, mode, cfl, nsteps);
    if (!c->use_graph || nsteps < 2) {
        for (int64_t s = 0; s < nsteps; ++s) {
            const int rc = enqueue_step(c, mode, cfl);
            if (rc) return rc;
        }
        return SWE_OK;
    }
    if (!c->graph || c->graph_mode != mode || c->graph_rows[0] != c->stage_rows[0] ||
        c->graph_rows[1] != c->stage_rows[1] || c->graph_rows[2] != c->stage_rows[2] ||
        c->graph_flat != (int)flat_bottom(c) ||
        c->graph_precision != c->precision || c->graph_stream != c->stream ||
        c->graph_scheme != scheme_key(c) ||
        (mode == 1 && c->graph_cfl != cfl)) {
        drop_graph(c);
        hipGraph_t g = nullptr;
        SWE_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = SWE_OK;
        for (int k = 0; k < kGraphSteps && rc == SWE_OK; ++k) rc = enqueue_step(c, mode, cfl);
        hipGraph_t captured = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &captured);
        if (rc) {
            if (captured) (void)hipGraphDestroy(captured);
            return rc;
        }
        SWE_CHECK(c, e);
        g = captured;
        const hipError_t ei = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SWE_CHECK(c, ei);
        c->graph_mode = mode;
        for (int i = 0; i < 3; ++i) c->graph_rows[i] = c->stage_rows[i];
        c->graph_flat = (int)flat_bottom(c);
        c->graph_precision = c->precision;
        c->graph_scheme = scheme_key(c);
        c->graph_stream = c->stream;
        c->graph_cfl = cfl;
    }
    int64_t s = 0;
    for (; s + kGraphSteps <= nsteps; s += kGraphSteps)
        SWE_CHECK(c, hipGraphLaunch(c->graph, c->stream));
    for (; s < nsteps; ++s) {
        const int rc = enqueue_step(c, mode, cfl);
        if (rc) return rc;
    }
    return SWE_OK;
}

// ---- scratch of the allocation-free adaptive step ------------------------------------
// Per-block maxima of the CFL wave-speed reduction (kNoAllocPartials doubles) followed by
// the dt-history counter, in a file-scope device array: each GPU the library runs on holds
// one copy (516 KiB) until the process ends.  Host code never passes the array's name to a
// kernel (in host code it is a stand-in, not a device address); it uses the address
// hipGetSymbolAddress returns for the current device, looked up once per device.
constexpr int kNoAllocPartials = 65536;
__device__ __align__(256) double scratch[kNoAllocPartials + 32];
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

int scratch_for_device(swe_hip_ctx *c, double **out) {
    if (c->device < 0 || c->device >= kMaxDevices)
        return fail(c, SWE_ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[c->device]) {
        void *ptr = nullptr;
        SWE_CHECK(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));  // device is current
        g_scratch_ptr[c->device] = ptr;
    }
    *out = static_cast<double *>(g_scratch_ptr[c->device]);
    return SWE_OK;
}

// Status words of a fused-step call of nsteps launches (fused_step_kernel): word 1 + (k & 1)
// of launch k holds bit s-1 for a non-finite stage s.  On a failure in step k the state the
// reference leaves is U^{k-1} (stage 1 or 2) or the stage-3 output of step k, i.e. the
// pointer set after k-1 or k swaps; the host made nsteps, so one more swap when the
// difference is odd.
int check_flag_fused(swe_hip_ctx *c, int64_t nsteps) {
    int w[swe::kStatusWords] = {};
    SWE_CHECK(c, hipMemcpyAsync(w, c->flag, sizeof(w), hipMemcpyDeviceToHost, c->stream));
    SWE_CHECK(c, hipStreamSynchronize(c->stream));
    c->last_status_step = w[0];
    const int bits = (w[1] | w[2]) & 7;
    c->last_status_stage = bits & 1 ? 1 : (bits & 2 ? 2 : (bits & 4 ? 3 : 0));
    if (c->last_status_stage == 0) return SWE_OK;
    const int64_t k = w[0], m = c->last_status_stage == 3 ? k : k - 1;
    if (((nsteps - m) & 1) != 0)
        for (int i = 0; i < 3; ++i) std::swap(c->U[i], c->Unext[i]);
    return fail(c, SWE_ERR_NONFINITE,
                "non-finite state detected in HIP solver at stage " +
                    std::to_string(c->last_status_stage) + " of step " + std::to_string(k) +
                    " of this call");
}

// Reads the run status words of the call that just ended.  A non-finite stage stopped the
// call there (the kernels' first-bad-stage stop): the error names the step (1-based,
// counted from this call's first) and the stage, and the arrays are as that stage left them.
int check_flag(swe_hip_ctx *c) {
    int w[swe::kStatusWords] = {};
    SWE_CHECK(c, hipMemcpyAsync(w, c->flag, sizeof(w), hipMemcpyDeviceToHost, c->stream));
    SWE_CHECK(c, hipStreamSynchronize(c->stream));
    c->last_status_step = w[0];
    c->last_status_stage = 0;
    for (int s = 1; s <= 3; ++s)
        if (w[s] & swe::kStatusBad) c->last_status_stage = s;
    if (c->last_status_stage == 0) return SWE_OK;
    if (w[0] == 0)  // a right-hand-side evaluation alone
        return fail(c, SWE_ERR_NONFINITE,
                    "non-finite state detected in HIP solver (right-hand side)");
    return fail(c, SWE_ERR_NONFINITE,
                "non-finite state detected in HIP solver at stage " +
                    std::to_string(c->last_status_stage) + " of step " + std::to_string(w[0]) +
                    " of this call");
}

int upload_common(swe_hip_ctx *c, const double *h, const double *hu, const double *hv,
                  const double *b, hipMemcpyKind kind) {
    const size_t bytes = c->n_total * sizeof(double);
    SWE_CHECK(c, hipMemcpyAsync(c->U[0], h, bytes, kind, c->stream));
    SWE_CHECK(c, hipMemcpyAsync(c->U[1], hu, bytes, kind, c->stream));
    SWE_CHECK(c, hipMemcpyAsync(c->U[2], hv, bytes, kind, c->stream));
    SWE_CHECK(c, hipMemcpyAsync(c->B, b, bytes, kind, c->stream));
    // B ghosts as at solver construction (omp_solver.hpp:119); state ghosts rebuilt too.
    SWE_CHECK(c, swe::launch_ghost_fill(c->bc, c->G, c->B, nullptr, nullptr, c->stream));
    SWE_CHECK(c, swe::launch_ghost_fill(c->bc, c->G, c->U[0], c->U[1], c->U[2], c->stream));
    // is the bottom +0.0 everywhere (ghosts included)?
    SWE_CHECK(c, swe::launch_fill_int(c->bz, 0, 1, c->stream));
    SWE_CHECK(c, swe::launch_nonzero_bits(c->B, (long)c->n_total, c->bz, c->stream));
    int nz = 1;
    SWE_CHECK(c, hipMemcpyAsync(&nz, c->bz, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SWE_CHECK(c, hipStreamSynchronize(c->stream));
    c->b_zero = nz == 0;
    return SWE_OK;
}

}  // namespace

extern "C" {

int swe_hip_abi_version(void) { return SWE_HIP_ABI_VERSION; }

const char *swe_hip_error_string(int code) {
    switch (code) {
        case SWE_OK: return "ok";
        case SWE_ERR_INVALID: return "invalid argument";
        case SWE_ERR_HIP: return "HIP runtime error";
        case SWE_ERR_NONFINITE: return "non-finite state detected";
        case SWE_ERR_CFL: return "time step exceeds the CFL limit";
        case SWE_ERR_OOM: return "device allocation failed";
        case SWE_ERR_NO_DEVICE: return "no HIP device";
        default: return "unknown error";
    }
}

const char *swe_hip_last_error(const swe_hip_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int swe_hip_last_status(const swe_hip_ctx *c, int64_t *step, int *stage) {
    if (!c) return SWE_ERR_INVALID;
    if (step) *step = c->last_status_step;
    if (stage) *stage = c->last_status_stage;
    return SWE_OK;
}

int swe_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int swe_hip_create(const swe_grid_desc *d, int device, swe_hip_ctx **out) {
    g_create_err.clear();
    auto reject = [](int code, const char *m) {
        g_create_err = m;
        return code;
    };
    if (!d || !out) return reject(SWE_ERR_INVALID, "swe_hip_create: null descriptor or output");
    *out = nullptr;
    if (d->ng < 2 || d->nx < d->ng || d->ny < d->ng || !(d->dx > 0.0) || !(d->dy > 0.0) ||
        !(d->g > 0.0) || (d->bc_kind != SWE_BC_REFLECTING && d->bc_kind != SWE_BC_TRANSMISSIVE))
        return reject(SWE_ERR_INVALID,
                      "swe_hip_create: need ng >= 2, nx and ny >= ng, dx, dy, g > 0 and a "
                      "known boundary kind");
    const long long nxt = (long long)d->nx + 2LL * d->ng, nyt = (long long)d->ny + 2LL * d->ng;
    if (nxt * nyt >= (1LL << 31))  // 32-bit cell indices
        return reject(SWE_ERR_INVALID, "swe_hip_create: (nx+2ng)*(ny+2ng) must be below 2^31");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return reject(SWE_ERR_NO_DEVICE, "swe_hip_create: no HIP device");
    if (device < 0 || device >= ndev)
        return reject(SWE_ERR_INVALID, "swe_hip_create: device index out of range");

    auto *c = new swe_hip_ctx();
    c->device = device;
    c->bc = d->bc_kind;
    c->G.nx = d->nx;
    c->G.ny = d->ny;
    c->G.ng = d->ng;
    c->G.stride = (int)nyt;
    c->G.dx = d->dx;
    c->G.dy = d->dy;
    c->G.inv_dx = 1.0 / d->dx;
    c->G.inv_dy = 1.0 / d->dy;
    c->G.g = d->g;
    c->G.half_g = 0.5 * d->g;
    c->G.split_ok = (d->g >= 0x1p-10 && d->g <= 0x1p10) ? 1 : 0;
    c->n_total = (size_t)(nxt * nyt);
    default_stage_rows(c->G, c->stage_rows);

    auto bad = [&](int code, const std::string &m) {
        swe_hip_destroy(c);
        g_create_err = "swe_hip_create: " + m + " failed";
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return bad(SWE_ERR_HIP, "hipSetDevice");
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess)
        return bad(SWE_ERR_HIP, "stream");
    c->stream = c->own_stream;
    const size_t bytes = c->n_total * sizeof(double);
    double **arrays[] = {&c->U[0], &c->U[1], &c->U[2], &c->U1[0], &c->U1[1],
                         &c->U1[2], &c->U2[0], &c->U2[1], &c->U2[2], &c->B};
    for (double **a : arrays)
        if (hipMalloc(a, bytes) != hipSuccess) return bad(SWE_ERR_OOM, "hipMalloc");
    if (hipMalloc(&c->scal, 4 * sizeof(double)) != hipSuccess) return bad(SWE_ERR_OOM, "scal");
    if (hipMalloc(&c->flag, sizeof(int) * swe::kStatusWords) != hipSuccess)
        return bad(SWE_ERR_OOM, "status words");
    if (hipMalloc(&c->counter, sizeof(long long)) != hipSuccess) return bad(SWE_ERR_OOM, "ctr");
    if (hipMalloc(&c->bz, sizeof(int)) != hipSuccess) return bad(SWE_ERR_OOM, "bottom word");
    for (double **a : arrays)
        if (swe::launch_fill_double(*a, 0.0, (long)c->n_total, c->stream) != hipSuccess)
            return bad(SWE_ERR_HIP, "fill");
    if (swe::launch_fill_double(c->scal, 0.0, 4, c->stream) != hipSuccess ||
        swe::launch_fill_int(c->flag, 0, swe::kStatusWords, c->stream) != hipSuccess)
        return bad(SWE_ERR_HIP, "fill");
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bad(SWE_ERR_HIP, "sync");
    *out = c;
    return SWE_OK;
}

void swe_hip_destroy(swe_hip_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (int i = 0; i < 3; ++i) {
        (void)hipFree(c->U[i]);
        (void)hipFree(c->U1[i]);
        (void)hipFree(c->U2[i]);
        (void)hipFree(c->Unext[i]);
    }
    (void)hipFree(c->B);
    (void)hipFree(c->scal);
    (void)hipFree(c->partial);
    (void)hipFree(c->flag);
    (void)hipFree(c->counter);
    (void)hipFree(c->bz);
    (void)hipFree(c->hist);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

int swe_hip_upload(swe_hip_ctx *c, const double *h, const double *hu, const double *hv,
                   const double *b) {
    if (!c || !h || !hu || !hv || !b) return SWE_ERR_INVALID;
    SWE_CHECK(c, hipSetDevice(c->device));
    return upload_common(c, h, hu, hv, b, hipMemcpyHostToDevice);
}

int swe_hip_upload_device(swe_hip_ctx *c, const double *h, const double *hu, const double *hv,
                          const double *b) {
    if (!c || !h || !hu || !hv || !b) return SWE_ERR_INVALID;
    SWE_CHECK(c, hipSetDevice(c->device));
    return upload_common(c, h, hu, hv, b, hipMemcpyDeviceToDevice);
}

static int download_common(swe_hip_ctx *c, double *h, double *hu, double *hv,
                           hipMemcpyKind kind) {
    const size_t bytes = c->n_total * sizeof(double);
    SWE_CHECK(c, swe::launch_ghost_fill(c->bc, c->G, c->U[0], c->U[1], c->U[2], c->stream));
    SWE_CHECK(c, hipMemcpyAsync(h, c->U[0], bytes, kind, c->stream));
    SWE_CHECK(c, hipMemcpyAsync(hu, c->U[1], bytes, kind, c->stream));
    SWE_CHECK(c, hipMemcpyAsync(hv, c->U[2], bytes, kind, c->stream));
    SWE_CHECK(c, hipStreamSynchronize(c->stream));
    return SWE_OK;
}

int swe_hip_download(swe_hip_ctx *c, double *h, double *hu, double *hv) {
    if (!c || !h || !hu || !hv) return SWE_ERR_INVALID;
    SWE_CHECK(c, hipSetDevice(c->device));
    return download_common(c, h, hu, hv, hipMemcpyDeviceToHost);
}

int swe_hip_download_device(swe_hip_ctx *c, double *h, double *hu, double *hv) {
    if (!c || !h || !hu || !hv) return SWE_ERR_INVALID;
    SWE_CHECK(c, hipSetDevice(c->device));
    return download_common(c, h, hu, hv, hipMemcpyDeviceToDevice);
}

int swe_hip_stable_dt(swe_hip_ctx *c, double cfl, double *dt_out, double *ms_out) {
    if (!c) return SWE_ERR_INVALID;
    SWE_CHECK(c, hipSetDevice(c->device));
    const int blocks = 1024;
    int rc = ensure_partial(c, blocks);
    if (rc) return rc;
    SWE_CHECK(c, swe::launch_speed_reduce(c->G, c->U[0], c->U[1], c->U[2], c->partial, blocks,
                                          c->stream));
    SWE_CHECK(c, swe::launch_finalize_dt(c->partial, blocks, cfl, c->G.dx, c->G.dy, c->scal,
                                         nullptr, nullptr, 0, nullptr, c->stream));
    double v[2];
    SWE_CHECK(c, hipMemcpyAsync(v, c->scal, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    SWE_CHECK(c, hipStreamSynchronize(c->stream));
    if (dt_out) *dt_out = v[0];
    if (ms_out) *ms_out = v[1];
    return SWE_OK;
}

int swe_hip_step(swe_hip_ctx *c, double dt, int64_t nsteps) {
    if (!c || nsteps < 0 || !(dt > 0.0)) return SWE_ERR_INVALID;
    SWE_CHECK(c, hipSetDevice(c->device));
    SWE_CHECK(c, swe::launch_set_scalar(c->scal, dt, c->stream));
    SWE_CHECK(c, swe::launch_fill_int(c->flag, 0, swe::kStatusWords, c->stream));
    const int rc = run_steps(c, 0, 0.0, nsteps);
    if (rc) return rc;
    return fused_mode(c) ? check_flag_fused(c, nsteps) : check_flag(c);
}

int swe_hip_step_adaptive(swe_hip_ctx *c, double cfl, int64_t nsteps, double *dt_hist) {
    if (!c || nsteps < 0 || !(cfl > 0.0)) return SWE_ERR_INVALID;
    if (c->scheme != SWE_SCHEME_OPTIMIZED)
        return fail(c, SWE_ERR_INVALID, "the serial scheme family runs with a fixed dt");
    SWE_CHECK(c, hipSetDevice(c->device));
    const dim3 grid = swe::stage_grid(c->G, c->stage_rows[2]);
    int need = (int)std::max<unsigned>(1024u, grid.x * grid.y);
    if (fused_mode(c)) need = std::max(need, swe::fused_step_blocks(c->G, c->fused_rows));
    int rc = ensure_partial(c, need);
    if (rc) return rc;
    if (c->hist_cap < nsteps + 1) {
        drop_graph(c);  // the graph bakes the history pointer
        if (c->hist) (void)hipFree(c->hist);
        c->hist = nullptr;
        c->hist_cap = 0;
        SWE_CHECK(c, hipMalloc(&c->hist, sizeof(double) * (size_t)(nsteps + 1)));
        c->hist_cap = nsteps + 1;
    }
    SWE_CHECK(c, swe::launch_fill_int(reinterpret_cast<int *>(c->counter), 0, 2, c->stream));
    SWE_CHECK(c, swe::launch_fill_int(c->flag, 0, swe::kStatusWords, c->stream));
    // dt of the first step from U^0
    SWE_CHECK(c, swe::launch_speed_reduce(c->G, c->U[0], c->U[1], c->U[2], c->partial, 1024,
                                          c->stream));
    SWE_CHECK(c, swe::launch_finalize_dt(c->partial, 1024, cfl, c->G.dx, c->G.dy, c->scal,
                                         c->hist, c->counter, c->hist_cap, c->flag, c->stream));
    rc = run_steps(c, 1, cfl, nsteps);
    if (rc) return rc;
    if (dt_hist && nsteps > 0)
        SWE_CHECK(c, hipMemcpyAsync(dt_hist, c->hist, sizeof(double) * (size_t)nsteps,
                                    hipMemcpyDeviceToHost, c->stream));
    return fused_mode(c) ? check_flag_fused(c, nsteps) : check_flag(c);
}

int swe_hip_step_adaptive_noalloc(swe_hip_ctx *c, double cfl, int64_t nsteps,
                                  double *dt_hist_device) {
    if (!c || nsteps < 0 || !(cfl > 0.0)) return SWE_ERR_INVALID;
    if (c->scheme != SWE_SCHEME_OPTIMIZED)
        return fail(c, SWE_ERR_INVALID, "the serial scheme family runs with a fixed dt");
    SWE_CHECK(c, hipSetDevice(c->device));
    const dim3 g2 = swe::stage_grid(c->G, c->stage_rows[2]);
    const long long blocks = (long long)g2.x * g2.y;
    if (blocks > kNoAllocPartials)
        return fail(c, SWE_ERR_INVALID,
                    "swe_hip_step_adaptive_noalloc: the stage launch has " +
                        std::to_string(blocks) + " blocks, more than the " +
                        std::to_string(kNoAllocPartials) +
                        " per-block maxima its fixed scratch holds (use a taller tile or "
                        "swe_hip_step_adaptive)");
    double *partial = nullptr;
    int rc = scratch_for_device(c, &partial);
    if (rc) return rc;
    long long *counter = reinterpret_cast<long long *>(partial + kNoAllocPartials);
    const long long hist_cap = dt_hist_device ? nsteps : 0;
    // the counter and the flag start at 0 (fill kernels, no memset node in a capture)
    SWE_CHECK(c, swe::launch_fill_int(reinterpret_cast<int *>(counter), 0, 2, c->stream));
    SWE_CHECK(c, swe::launch_fill_int(c->flag, 0, swe::kStatusWords, c->stream));
    // dt of the first step from U^0
    SWE_CHECK(c, swe::launch_speed_reduce(c->G, c->U[0], c->U[1], c->U[2], partial, 1024,
                                          c->stream));
    SWE_CHECK(c, swe::launch_finalize_dt(partial, 1024, cfl, c->G.dx, c->G.dy, c->scal,
                                         dt_hist_device, counter, hist_cap, c->flag, c->stream));
    for (int64_t s = 0; s < nsteps; ++s) {
        for (int stage = 1; stage <= 3; ++stage) {
            StageArgs a = stage_args(c, stage);
            a.partial = partial;
            dim3 grid;
            SWE_CHECK(c, launch_ctx_stage(c, stage, stage == 3, a, &grid));
            if (stage == 3)
                SWE_CHECK(c, swe::launch_finalize_dt(partial, (int)(grid.x * grid.y), cfl,
                                                     c->G.dx, c->G.dy, c->scal, dt_hist_device,
                                                     counter, hist_cap, c->flag, c->stream));
        }
    }
    return check_flag(c);
}

int swe_hip_rhs(swe_hip_ctx *c, double *rh, double *rhu, double *rhv) {
    if (!c || !rh || !rhu || !rhv) return SWE_ERR_INVALID;
    SWE_CHECK(c, hipSetDevice(c->device));
    // rhs goes to the U1 buffers (zeroed first so ghosts read back as 0)
    for (int i = 0; i < 3; ++i)
        SWE_CHECK(c, swe::launch_fill_double(c->U1[i], 0.0, (long)c->n_total, c->stream));
    StageArgs a = stage_args(c, 0);
    SWE_CHECK(c, launch_ctx_stage(c, 0, false, a, nullptr));
    const size_t bytes = c->n_total * sizeof(double);
    SWE_CHECK(c, hipMemcpyAsync(rh, c->U1[0], bytes, hipMemcpyDeviceToHost, c->stream));
    SWE_CHECK(c, hipMemcpyAsync(rhu, c->U1[1], bytes, hipMemcpyDeviceToHost, c->stream));
    SWE_CHECK(c, hipMemcpyAsync(rhv, c->U1[2], bytes, hipMemcpyDeviceToHost, c->stream));
    SWE_CHECK(c, hipStreamSynchronize(c->stream));
    return SWE_OK;
}

int swe_hip_set_tile_rows(swe_hip_ctx *c, int tile_rows) {
    if (!c || tile_rows < 1) return SWE_ERR_INVALID;
    // one segment covers every row; larger values would only overflow r0 + tile_rows
    for (int &r : c->stage_rows) r = std::min(tile_rows, c->G.nx);
    return SWE_OK;
}

int swe_hip_get_tile_rows(const swe_hip_ctx *c) { return c ? c->stage_rows[0] : -1; }

int swe_hip_set_flat_bottom(swe_hip_ctx *c, int mode) {
    if (!c || (mode != 0 && mode != 1)) return SWE_ERR_INVALID;
    c->flat_bottom_mode = mode;
    return SWE_OK;
}

int swe_hip_get_flat_bottom(const swe_hip_ctx *c) {
    if (!c) return -1;
    return flat_bottom(c) ? 1 : 0;
}

int swe_hip_set_stage_tile_rows(swe_hip_ctx *c, int stage, int tile_rows) {
    if (!c || stage < 1 || stage > 3 || tile_rows < 1) return SWE_ERR_INVALID;
    c->stage_rows[stage - 1] = std::min(tile_rows, c->G.nx);
    return SWE_OK;
}

int swe_hip_get_stage_tile_rows(const swe_hip_ctx *c, int stage) {
    return (c && stage >= 1 && stage <= 3) ? c->stage_rows[stage - 1] : -1;
}

int swe_hip_set_precision(swe_hip_ctx *c, int mode) {
    if (!c || (mode != SWE_PRECISION_EXACT && mode != SWE_PRECISION_FAST)) return SWE_ERR_INVALID;
    c->precision = mode;
    return SWE_OK;
}

int swe_hip_get_precision(const swe_hip_ctx *c) { return c ? c->precision : -1; }

int swe_hip_set_stream(swe_hip_ctx *c, void *stream) {
    if (!c) return SWE_ERR_INVALID;
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    return SWE_OK;
}

int swe_hip_set_scheme(swe_hip_ctx *c, int scheme, int recon, int riemann) {
    if (!c) return SWE_ERR_INVALID;
    if (scheme == SWE_SCHEME_OPTIMIZED) {
        c->scheme = scheme;
        c->recon = SWE_RECON_MUSCL;
        c->riemann = SWE_RIEMANN_HLL;
        return SWE_OK;
    }
    if (scheme != SWE_SCHEME_SERIAL && scheme != SWE_SCHEME_SERIAL_PLAIN)
        return fail(c, SWE_ERR_INVALID, "unknown scheme");
    if (c->bc != SWE_BC_REFLECTING)
        return fail(c, SWE_ERR_INVALID, "the serial scheme family has reflecting walls only");
    if (c->G.ng != 2)
        return fail(c, SWE_ERR_INVALID, "the serial scheme family needs ghost width ng = 2");
    if (!swe::modular_supported(scheme == SWE_SCHEME_SERIAL, recon, riemann))
        return fail(c, SWE_ERR_INVALID, "Unsupported serial solver combination");
    c->scheme = scheme;
    c->recon = recon;
    c->riemann = riemann;
    return SWE_OK;
}

int swe_hip_set_step_fusion(swe_hip_ctx *c, int rows) {
    if (!c || rows < 0) return SWE_ERR_INVALID;
    SWE_CHECK(c, hipSetDevice(c->device));
    if (rows == 0) {
        c->fused_rows = 0;
        return SWE_OK;
    }
    if (c->G.ng != 2)
        return fail(c, SWE_ERR_INVALID, "fused steps need ghost width ng = 2");
    if (!c->Unext[0]) {
        const size_t bytes = c->n_total * sizeof(double);
        for (int i = 0; i < 3; ++i) {
            if (hipMalloc(&c->Unext[i], bytes) != hipSuccess) {
                for (int k = 0; k <= i; ++k) {
                    (void)hipFree(c->Unext[k]);
                    c->Unext[k] = nullptr;
                }
                return fail(c, SWE_ERR_OOM, "fused steps: second state buffer");
            }
            SWE_CHECK(c, swe::launch_fill_double(c->Unext[i], 0.0, (long)c->n_total, c->stream));
        }
    }
    c->fused_rows = std::max(4, std::min(rows, c->G.nx));
    return SWE_OK;
}

int swe_hip_get_step_fusion(const swe_hip_ctx *c) { return c ? c->fused_rows : -1; }

int swe_hip_get_scheme(const swe_hip_ctx *c, int *scheme, int *recon, int *riemann) {
    if (!c) return SWE_ERR_INVALID;
    if (scheme) *scheme = c->scheme;
    if (recon) *recon = c->recon;
    if (riemann) *riemann = c->riemann;
    return SWE_OK;
}

int swe_hip_set_graph(swe_hip_ctx *c, int enabled) {
    if (!c) return SWE_ERR_INVALID;
    c->use_graph = enabled != 0;
    if (!c->use_graph) drop_graph(c);
    return SWE_OK;
}

int swe_hip_time_steps(swe_hip_ctx *c, double dt, int64_t nsteps, double *stage_ms,
                       double *total_ms) {
    if (!c || nsteps < 1 || !(dt > 0.0)) return SWE_ERR_INVALID;
    SWE_CHECK(c, hipSetDevice(c->device));
    SWE_CHECK(c, swe::launch_set_scalar(c->scal, dt, c->stream));
    SWE_CHECK(c, swe::launch_fill_int(c->flag, 0, swe::kStatusWords, c->stream));
    swe::EventSet ev((size_t)nsteps * 3 + 1);
    SWE_CHECK(c, ev.status);
    int rc = SWE_OK;
    SWE_CHECK(c, hipEventRecord(ev[0], c->stream));
    for (int64_t s = 0; s < nsteps && rc == SWE_OK; ++s) {
        for (int stage = 1; stage <= 3; ++stage) {
            StageArgs a = stage_args(c, stage);
            const hipError_t e = launch_ctx_stage(c, stage, false, a, nullptr);
            if (e != hipSuccess) {
                rc = fail(c, SWE_ERR_HIP, std::string("launch_stage: ") + hipGetErrorString(e));
                break;
            }
            SWE_CHECK(c, hipEventRecord(ev[(size_t)s * 3 + stage], c->stream));
        }
    }
    SWE_CHECK(c, hipStreamSynchronize(c->stream));
    if (rc == SWE_OK) {
        double acc[3] = {0, 0, 0};
        for (int64_t s = 0; s < nsteps; ++s)
            for (int stage = 1; stage <= 3; ++stage) {
                float ms = 0.f;
                SWE_CHECK(c, hipEventElapsedTime(&ms, ev[(size_t)s * 3 + stage - 1],
                                                 ev[(size_t)s * 3 + stage]));
                acc[stage - 1] += ms;
            }
        float tot = 0.f;
        SWE_CHECK(c, hipEventElapsedTime(&tot, ev[0], ev[(size_t)nsteps * 3]));
        if (stage_ms)
            for (int i = 0; i < 3; ++i) stage_ms[i] = acc[i] / (double)nsteps;
        if (total_ms) *total_ms = tot;
    }
    if (rc) return rc;
    return check_flag(c);
}

int swe_hip_sync(swe_hip_ctx *c) {
    if (!c) return SWE_ERR_INVALID;
    SWE_CHECK(c, hipStreamSynchronize(c->stream));
    return SWE_OK;
}

}  // extern "C"
