// swe_adjoint_api.hip -- discrete adjoint driver: binomial checkpointing over the fused
// stage kernels, and the adjoint entry points of include/swe_hip.h.
//
// Checkpoint schedule.  Reversing n steps whose start state sits in a stored slot, with f
// more free slots (s = f + 1 stored states in all), costs at least r*n - C(s+r, r-1)
// single-step advances, r the least integer with C(s+r, s) >= n (Griewank's binomial bound,
// A. Griewank and A. Walther, "Algorithm 799: revolve", ACM TOMS 26(1), 2000).  The
// recursion below attains it: advance k steps, store that state, reverse the right part
// with f-1 free slots, then the left part with f; any k in
//     [max(C(s+r-2, s), n - C(s-1+r, s-1)), min(C(s+r-1, s), n - C(s-1+r-1, s-1))]
// is optimal, and the lower end is taken.  With f = 0 every step is reached by advancing
// from the stored start.  Slots are used last-in first-out.
//
// Per reversed step i (state U^i = "base", dt_i frozen): recompute U1 and U2 with the
// forward stage kernels, then run the adjoint stage kernels 3, 2, 1.  The floor masks of
// the stage outputs travel between the adjoint kernels (each writes the mask of its own
// input), so no stage output is ever recomputed for them.
//
// Stage records.  The other storage plan keeps "records" -- U^p, U1 and U2 of a step, the
// three inputs its adjoint stages read -- for the last K steps straight from the forward
// pass, and a checkpoint of U^p every m steps before them.  The backward reverses the
// recorded steps as they are, then, segment by segment from the last, recomputes the m
// steps of a segment from its checkpoint into the record buffer (stage 3 of the segment's
// last step is not needed) and reverses them.  Every step before the recorded tail is
// recomputed exactly once.  With two segments or more, the record buffer holds two segment
// buffers and a second stream recomputes the next segment into one while the first stream
// reverses the other (events order them), so the forward kernels fill the CUs the adjoint
// kernels leave idle.  K = nsteps is "store every step".  plan_records picks m and K for a
// memory budget: the largest K (fewest recomputed steps) whose records, max(K, m or 2m) x
// 9 arrays, and checkpoints, ceil((nsteps - K) / m) x 3 arrays, fit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/swe_hip.h"
#include "swe_ctx.h"
#include "swe_internal.h"

namespace {

enum { kAdvance = 0, kAdjoint = 1 };

// kAdvance: state a (in slot src, -1 = work buffer) -> state b (into slot dst, -1 = work).
// kAdjoint: reverse step a whose start state is in slot src (-1 = work buffer).
struct Action {
    int kind;
    int64_t a, b;
    int src, dst;
};

// C(s + r, r) saturated above `cap`.
int64_t beta(int64_t s, int64_t r, int64_t cap) {
    if (r < 0 || s < 0) return 0;
    double b = 1.0;
    for (int64_t i = 1; i <= r; ++i) {
        b = b * (double)(s + i) / (double)i;
        if (b > (double)cap) return cap + 1;
    }
    return (int64_t)(b + 0.5);
}

int64_t repetitions(int64_t n, int64_t s) {
    int64_t r = 0;
    while (beta(s, r, n) < n) ++r;
    return r;
}

std::vector<Action> make_schedule(int64_t nsteps, int32_t nsnap) {
    struct Task {
        int64_t a, b;
        int f, slot;
    };
    std::vector<Action> out;
    std::vector<Task> stack{{0, nsteps, nsnap - 1, 0}};
    while (!stack.empty()) {
        const Task t = stack.back();
        stack.pop_back();
        const int64_t n = t.b - t.a;
        if (n == 1) {
            out.push_back({kAdjoint, t.a, t.b, t.slot, -1});
            continue;
        }
        if (t.f == 0) {
            for (int64_t i = t.b - 1; i >= t.a; --i) {
                if (i > t.a) out.push_back({kAdvance, t.a, i, t.slot, -1});
                out.push_back({kAdjoint, i, i + 1, i > t.a ? -1 : t.slot, -1});
            }
            continue;
        }
        const int64_t s = t.f + 1;
        const int64_t r = repetitions(n, s);
        const int64_t k = std::max<int64_t>({1, beta(s, r - 2, n), n - beta(s - 1, r, n)});
        const int64_t m = t.a + k;
        const int slot_m = nsnap - t.f;
        out.push_back({kAdvance, t.a, m, t.slot, slot_m});
        stack.push_back({t.a, m, t.f, t.slot});      // reversed after the right part
        stack.push_back({m, t.b, t.f - 1, slot_m});  // next
    }
    return out;
}

// Records plan (see the header): K recorded steps, a checkpoint every m steps before them.
struct RecordPlan {
    int64_t K = 0, m = 1, nck = 0;
    // record slots: the tail's K, and two segment buffers when there are two segments or
    // more (one recomputed while the other is reversed)
    int64_t slots() const { return std::max(K, nck >= 2 ? 2 * m : m); }
    int64_t arrays() const { return 9 * slots() + 3 * nck; }
};

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// The plan with the most recorded steps whose storage fits `arrays_budget` arrays (ties:
// the smaller storage); K = 0 when none fits.
RecordPlan plan_records(int64_t nsteps, int64_t arrays_budget) {
    RecordPlan best;
    if (9 * nsteps <= arrays_budget) {
        best.K = nsteps;
        return best;
    }
    for (int64_t m = 1; m <= nsteps; ++m) {
        auto need = [&](int64_t K) {
            const int64_t nck = ceil_div(nsteps - K, m);
            return 9 * std::max(K, nck >= 2 ? 2 * m : m) + 3 * nck;
        };
        // need() does not rise with K below kt = min(2m, nsteps - m) (the segment buffers
        // dominate the records while fewer steps are recorded) and does not fall above it
        int64_t kt = std::min(2 * m, nsteps - m);
        kt = std::max<int64_t>(1, std::min(kt, nsteps - 1));
        if (need(kt) > arrays_budget) continue;
        int64_t lo = kt, hi = nsteps - 1;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo + 1) / 2;
            if (need(mid) <= arrays_budget)
                lo = mid;
            else
                hi = mid - 1;
        }
        RecordPlan cand{lo, m, ceil_div(nsteps - lo, m)};
        if (cand.K > best.K || (cand.K == best.K && cand.arrays() < best.arrays())) best = cand;
    }
    return best;
}

}  // namespace

struct swe_hip_adj {
    swe_hip_ctx *ctx = nullptr;
    int64_t nsteps = 0;
    int32_t nsnap = 0;
    size_t n_total = 0;
    double *slots = nullptr;  // binomial: nsnap states of 3 arrays; records plan: the
                              // ckpt states (3 arrays each)
    bool records = false;     // records plan (stage records; binomial otherwise)
    int64_t rec_steps = 0;    // records plan: K recorded steps (the last K of the run)
    int64_t ckpt_every = 1;   // records plan: checkpoint interval m before them
    int64_t nck = 0;          // records plan: ceil((nsteps - K) / m) checkpoints
    int64_t rec_slots = 0;    // records plan: record slots (RecordPlan::slots())
    double *rec = nullptr;    // records plan: RecordPlan::slots() records of 9 arrays
                              // (U^p, U1, U2)
    hipStream_t stream2 = nullptr;  // the recompute stream (records plan with two segments
                                    // or more; binomial schedule with overlap buffers)
    double *ovl[9] = {};  // binomial overlap: a second work state (0..2) and a second pair
                          // of stage states (U1: 3..5, U2: 6..8); null without overlap
    int per_slot = 3;
    double *W[3] = {nullptr, nullptr, nullptr};
    double *lam[4][3] = {};   // step adjoint, dL/dU2, dL/dU1, U^n accumulator
    double *lam_b = nullptr;
    unsigned char *mask[2] = {nullptr, nullptr};
    double *dts = nullptr;
    double *partial = nullptr;  // 1024 per-block maxima (adaptive dt)
    double *xf = nullptr;       // two-kernel path: x-face outputs (8 per face and column)
    unsigned char *xsel = nullptr;  // two-kernel path: minmod selections of the x slopes
    long long *counter = nullptr;
    std::vector<Action> acts;
    size_t first_adj = 0;
    int block_waves = 4;  // fused kernel: waves per block (0 selects the two-kernel path)
    int fused_rows = 4;   // fused kernel: at most this many rows per wave
    int tile_rows = 4;    // two-kernel path: x faces per wave
    int y_rows = 4;       // two-kernel path: rows per wave of the y-face kernel
    int state = 0;  // 0 created, 1 forward done (checkpoints valid), 2 consumed
    int parity = 0;
    double fwd_ms = 0.0, bwd_ms = 0.0;
    std::string err;
};

namespace {

int afail(swe_hip_adj *a, int code, const std::string &msg) {
    if (a) a->err = msg;
    return code;
}

#define ADJ_CHECK(adj, call)                                                                \
    do {                                                                                    \
        const hipError_t e__ = (call);                                                      \
        if (e__ != hipSuccess)                                                              \
            return afail(adj, SWE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e__)); \
    } while (0)

double *slot_ptr(swe_hip_adj *a, int slot, int comp) {
    return a->slots + ((size_t)slot * a->per_slot + comp) * a->n_total;
}

struct State {
    double *p[3];
};

State state_of(swe_hip_adj *a, int slot) {
    if (slot < 0) return State{{a->W[0], a->W[1], a->W[2]}};
    return State{{slot_ptr(a, slot, 0), slot_ptr(a, slot, 1), slot_ptr(a, slot, 2)}};
}

// rows: tile height of the launch (0: the context's stage-1 height).
swe::StageArgs fwd_args(swe_hip_adj *a, const State &in, const State &u0, const State &out,
                        const double *dt, int rows = 0) {
    swe_hip_ctx *c = a->ctx;
    swe::StageArgs s{};
    s.in_h = in.p[0];
    s.in_hu = in.p[1];
    s.in_hv = in.p[2];
    s.u0_h = u0.p[0];
    s.u0_hu = u0.p[1];
    s.u0_hv = u0.p[2];
    s.B = c->B;
    s.out_h = out.p[0];
    s.out_hu = out.p[1];
    s.out_hv = out.p[2];
    s.dt = dt;
    s.partial = nullptr;
    s.flag = c->flag;
    s.tile_rows = rows > 0 ? std::min(rows, c->G.nx) : c->stage_rows[0];
    s.nstrips = swe::strips_for(c->G.ny);
    return s;
}

// Tile height of the segment recompute that runs beside the reversal on the second stream.
// The default heights fill the GPU when a grid steps on its own (4 rows at 1024^2), at the
// price of a 4-row prologue per tile; beside the adjoint kernels, which hold most wave
// slots, fewer and taller tiles cost less: at 1024^2 the records plan's backward took
// 560 / 574 / 575 / 596 ms with 8 / 4 / 16 / 32-row tiles (interleaved runs, one box).
int recompute_rows(const swe_hip_ctx *c) { return std::max(8, c->stage_rows[0]); }

// The context's stage buffers (stage states of a step that is not recorded).
State scratch_stage(swe_hip_adj *a, int which) {
    double *const *b = which == 1 ? a->ctx->U1 : a->ctx->U2;
    return State{{b[0], b[1], b[2]}};
}

// Record r of the records plan: which = 0 (U^p), 1 (U1), 2 (U2).
State rec_state(swe_hip_adj *a, int64_t r, int which) {
    double *base = a->rec + ((size_t)r * 9 + 3 * (size_t)which) * a->n_total;
    return State{{base, base + a->n_total, base + 2 * a->n_total}};
}

// Stages 1 and 2 of step p from `base` into u1, u2 (on `st`; default the context's stream;
// rows: tile height, 0 for the context's).
int stages_12(swe_hip_adj *a, const State &base, const State &u1, const State &u2, int64_t p,
              hipStream_t st = nullptr, int rows = 0) {
    swe_hip_ctx *c = a->ctx;
    if (!st) st = c->stream;
    ADJ_CHECK(a, swe::launch_stage(1, c->bc, false, false, c->G,
                                   fwd_args(a, base, base, u1, a->dts + p, rows), st, nullptr));
    ADJ_CHECK(a, swe::launch_stage(2, c->bc, false, false, c->G,
                                   fwd_args(a, u1, base, u2, a->dts + p, rows), st, nullptr));
    return SWE_OK;
}

// One forward step p: in -> out (out may alias in) through the stage states u1, u2.  With
// cfl > 0, dt_p is first computed from `in` (compute_stable_dt) and recorded in dts[p].
int forward_step_via(swe_hip_adj *a, const State &in, const State &u1, const State &u2,
                     const State &out, int64_t p, double cfl) {
    swe_hip_ctx *c = a->ctx;
    if (cfl > 0.0) {
        ADJ_CHECK(a, swe::launch_speed_reduce(c->G, in.p[0], in.p[1], in.p[2], a->partial, 1024,
                                              c->stream));
        ADJ_CHECK(a, swe::launch_finalize_dt(a->partial, 1024, cfl, c->G.dx, c->G.dy, c->scal,
                                             a->dts, a->counter, a->nsteps, c->flag,
                                             c->stream));
    }
    const int rc = stages_12(a, in, u1, u2, p);
    if (rc) return rc;
    ADJ_CHECK(a, swe::launch_stage(3, c->bc, false, false, c->G,
                                   fwd_args(a, u2, in, out, a->dts + p), c->stream, nullptr));
    return SWE_OK;
}

int forward_step(swe_hip_adj *a, const State &in, const State &out, int64_t p, double cfl) {
    return forward_step_via(a, in, scratch_stage(a, 1), scratch_stage(a, 2), out, p, cfl);
}

int advance(swe_hip_adj *a, const Action &act, double cfl) {
    const State src = state_of(a, act.src), work = state_of(a, -1);
    for (int64_t p = act.a; p < act.b; ++p) {
        const State in = p == act.a ? src : work;
        const State out = (p == act.b - 1 && act.dst >= 0) ? state_of(a, act.dst) : work;
        const int rc = forward_step(a, in, out, p, cfl);
        if (rc) return rc;
    }
    return SWE_OK;
}

swe::AdjArgs adj_args(swe_hip_adj *a, const State &v, int lam_in, int lam_out,
                      const unsigned char *mask_out, unsigned char *mask_in, int64_t p) {
    swe_hip_ctx *c = a->ctx;
    swe::AdjArgs r{};
    r.v_h = v.p[0];
    r.v_hu = v.p[1];
    r.v_hv = v.p[2];
    r.B = c->B;
    r.lam_h = a->lam[lam_in][0];
    r.lam_hu = a->lam[lam_in][1];
    r.lam_hv = a->lam[lam_in][2];
    r.mask_out = mask_out;
    r.out_h = a->lam[lam_out][0];
    r.out_hu = a->lam[lam_out][1];
    r.out_hv = a->lam[lam_out][2];
    r.e_h = a->lam[3][0];
    r.e_hu = a->lam[3][1];
    r.e_hv = a->lam[3][2];
    r.lam_b = a->lam_b;
    r.mask_in = mask_in;
    r.dt = a->dts + p;
    r.block_waves = a->block_waves;
    r.tile_rows = a->block_waves ? a->fused_rows : a->tile_rows;
    r.y_rows = a->y_rows;
    r.nstrips = swe::adj_strips_for(c->G.ny);
    r.xf = a->xf;
    r.xsel = a->xsel;
    return r;
}

// Reverse step p given its stage inputs U^p (base), U1, U2: lam[0] = dL/dU^{p+1} ->
// dL/dU^p.
int reverse_with(swe_hip_adj *a, const State &base, const State &u1, const State &u2,
                 int64_t p) {
    swe_hip_ctx *c = a->ctx;
    unsigned char *m0 = a->mask[a->parity], *m1 = a->mask[1 - a->parity];
    // stage 3: V = U2, output U^{p+1} (mask m0)
    ADJ_CHECK(a, swe::launch_adjoint_stage(3, c->bc, c->G, adj_args(a, u2, 0, 1, m0, m1, p),
                                           c->stream));
    // stage 2: V = U1, output U2 (mask m1, written by stage 3)
    ADJ_CHECK(a, swe::launch_adjoint_stage(2, c->bc, c->G, adj_args(a, u1, 1, 2, m1, m0, p),
                                           c->stream));
    // stage 1: V = U^p, output U1 (mask m0, written by stage 2); writes the mask of U^p
    ADJ_CHECK(a, swe::launch_adjoint_stage(1, c->bc, c->G, adj_args(a, base, 2, 0, m0, m1, p),
                                           c->stream));
    a->parity = 1 - a->parity;
    return SWE_OK;
}

// Binomial schedule: reverse step p from its start state in `src` (U1, U2 recomputed).
int reverse_step(swe_hip_adj *a, int src, int64_t p) {
    const State base = state_of(a, src);
    const State u1 = scratch_stage(a, 1), u2 = scratch_stage(a, 2);
    const int rc = stages_12(a, base, u1, u2, p);
    if (rc) return rc;
    return reverse_with(a, base, u1, u2, p);
}

// Records plan, forward: steps before the recorded tail keep a checkpoint every m steps
// (stage states in the scratch buffers); the tail's steps write their records.
int forward_records(swe_hip_adj *a, double cfl) {
    swe_hip_ctx *c = a->ctx;
    const int64_t N = a->nsteps, T = N - a->rec_steps, m = a->ckpt_every;
    const State work = state_of(a, -1), cu{{c->U[0], c->U[1], c->U[2]}};
    State in = T > 0 ? state_of(a, 0) : rec_state(a, 0, 0);
    const size_t bytes = a->n_total * sizeof(double);
    for (int k = 0; k < 3; ++k)
        ADJ_CHECK(a, hipMemcpyAsync(in.p[k], c->U[k], bytes, hipMemcpyDeviceToDevice,
                                    c->stream));
    for (int64_t p = 0; p < N; ++p) {
        const int64_t q = p + 1;
        State u1, u2, out;
        if (p < T) {
            u1 = scratch_stage(a, 1);
            u2 = scratch_stage(a, 2);
            out = q == T ? rec_state(a, 0, 0) : (q % m == 0 ? state_of(a, (int)(q / m)) : work);
        } else {
            u1 = rec_state(a, p - T, 1);
            u2 = rec_state(a, p - T, 2);
            out = q < N ? rec_state(a, q - T, 0) : cu;
        }
        const int rc = forward_step_via(a, in, u1, u2, out, p, cfl);
        if (rc) return rc;
        in = out;
    }
    return SWE_OK;
}

// Records plan, backward: the recorded tail as stored, then each checkpoint segment
// (last first) recomputed into a segment buffer of the record buffer and reversed.  With
// two segments or more the recompute runs on stream2, one segment ahead of the reversal,
// alternating between two segment buffers at the top of the record buffer (slots
// [R - m, R) and [R - 2m, R - m), R = its slot count).  The tail is reversed from its top
// slot down, so the first two recomputes start as soon as the tail's reversal has left
// their buffer, while the rest of the tail is still being reversed.
int backward_records(swe_hip_adj *a) {
    swe_hip_ctx *c = a->ctx;
    const int64_t N = a->nsteps, T = N - a->rec_steps, m = a->ckpt_every, K = a->rec_steps;
    const int64_t R = a->rec_slots;
    const bool overlap = a->nck >= 2 && a->stream2;
    // events: 0/1 buffer b left by the tail, 2/3 buffer b recomputed, 4/5 buffer b
    // reversed, 6 stream2 done
    swe::EventSet ev(overlap ? 7 : 0);
    if (overlap) ADJ_CHECK(a, ev.status);
    auto off = [&](int b) { return overlap ? R - (int64_t)(b + 1) * m : 0; };
    for (int b = 0; b < 2 && overlap; ++b)
        if (off(b) >= K) ADJ_CHECK(a, hipEventRecord(ev[b], c->stream));  // never used by the tail
    for (int64_t p = N - 1; p >= T; --p) {
        const int rc = reverse_with(a, rec_state(a, p - T, 0), rec_state(a, p - T, 1),
                                    rec_state(a, p - T, 2), p);
        if (rc) return rc;
        for (int b = 0; b < 2 && overlap; ++b)
            if (p - T == std::max<int64_t>(off(b), 0) && off(b) < K)
                ADJ_CHECK(a, hipEventRecord(ev[b], c->stream));
    }
    // segment seg: steps [s0, s1), records from slot o
    auto seg_range = [&](int64_t seg, int64_t &s0, int64_t &s1) {
        s0 = seg * m;
        s1 = std::min(s0 + m, T);
    };
    auto recompute = [&](int64_t seg, int64_t o, hipStream_t st) -> int {
        int64_t s0, s1;
        seg_range(seg, s0, s1);
        const State ck = state_of(a, (int)seg);
        const int rows = st != c->stream ? recompute_rows(c) : 0;
        for (int64_t p = s0; p < s1; ++p) {
            const State base = p == s0 ? ck : rec_state(a, o + p - s0, 0);
            const State u1 = rec_state(a, o + p - s0, 1), u2 = rec_state(a, o + p - s0, 2);
            const int rc = stages_12(a, base, u1, u2, p, st, rows);
            if (rc) return rc;
            if (p + 1 < s1)  // the segment's last step needs only its stage states
                ADJ_CHECK(a, swe::launch_stage(3, c->bc, false, false, c->G,
                                               fwd_args(a, u2, base,
                                                        rec_state(a, o + p + 1 - s0, 0),
                                                        a->dts + p, rows),
                                               st, nullptr));
        }
        return SWE_OK;
    };
    auto reverse = [&](int64_t seg, int64_t o) -> int {
        int64_t s0, s1;
        seg_range(seg, s0, s1);
        const State ck = state_of(a, (int)seg);
        for (int64_t p = s1 - 1; p >= s0; --p) {
            const State base = p == s0 ? ck : rec_state(a, o + p - s0, 0);
            const int rc = reverse_with(a, base, rec_state(a, o + p - s0, 1),
                                        rec_state(a, o + p - s0, 2), p);
            if (rc) return rc;
        }
        return SWE_OK;
    };
    if (!overlap) {
        for (int64_t seg = a->nck - 1; seg >= 0; --seg) {
            int rc = recompute(seg, 0, c->stream);
            if (!rc) rc = reverse(seg, 0);
            if (rc) return rc;
        }
        return SWE_OK;
    }
    hipStream_t s2 = a->stream2;
    const int64_t S = a->nck;
    ADJ_CHECK(a, hipStreamWaitEvent(s2, ev[0], 0));
    int rc = recompute(S - 1, off(0), s2);
    if (rc) return rc;
    ADJ_CHECK(a, hipEventRecord(ev[2], s2));
    for (int64_t seg = S - 1; seg >= 0; --seg) {
        const int b = (int)((S - 1 - seg) % 2);
        if (seg >= 1) {
            const int nb = 1 - b;
            // buffer nb: first use waits for the tail to leave it, later ones for the
            // reversal of the segment it last held (seg + 1)
            ADJ_CHECK(a, hipStreamWaitEvent(s2, seg + 1 <= S - 1 ? ev[4 + nb] : ev[nb], 0));
            rc = recompute(seg - 1, off(nb), s2);
            if (rc) return rc;
            ADJ_CHECK(a, hipEventRecord(ev[2 + nb], s2));
        }
        ADJ_CHECK(a, hipStreamWaitEvent(c->stream, ev[2 + b], 0));
        rc = reverse(seg, off(b));
        if (rc) return rc;
        ADJ_CHECK(a, hipEventRecord(ev[4 + b], c->stream));
    }
    ADJ_CHECK(a, hipEventRecord(ev[6], s2));
    ADJ_CHECK(a, hipStreamWaitEvent(c->stream, ev[6], 0));
    // the events are released on return: drain both streams' use of them first
    ADJ_CHECK(a, hipStreamSynchronize(c->stream));
    return SWE_OK;
}

// Binomial schedule, backward with the recomputation on stream2: the advances and the
// stage states of each reversed step are issued there, one reversal ahead of the adjoint
// kernels on the main stream.  The stage states alternate between two pairs (the
// context's U1/U2 and a second pair), the advances' work state between two buffers:
//   ready[k]  (stream2)  pair k holds the stage states of the reversal that reads it next;
//   done[k]   (main)     the last reversal issued with pair k has finished;
// each work buffer and checkpoint slot remembers the pair of the last reversal that read
// it, and stream2 waits for that reversal before it overwrites it (an event waited on is
// its latest record, a reversal at least that late, already issued after its own ready:
// no cycle).  The adjoint kernels read the same states and run in the same order as on
// one stream, so the gradient is the same bits.
int backward_binomial_overlap(swe_hip_adj *a) {
    swe_hip_ctx *c = a->ctx;
    hipStream_t s2 = a->stream2;
    swe::EventSet ev(5);  // 0/1 ready[k], 2/3 done[k], 4 stream hand-over
    ADJ_CHECK(a, ev.status);
    const int rows = recompute_rows(c);
    const State pair[2][2] = {{scratch_stage(a, 1), scratch_stage(a, 2)},
                              {State{{a->ovl[3], a->ovl[4], a->ovl[5]}},
                               State{{a->ovl[6], a->ovl[7], a->ovl[8]}}}};
    const State work[2] = {state_of(a, -1), State{{a->ovl[0], a->ovl[1], a->ovl[2]}}};
    std::vector<int> slot_reader((size_t)std::max(a->nsnap, 1), -1);
    int work_reader[2] = {-1, -1};
    bool issued[2] = {false, false};  // a reversal with pair k has been issued
    int last_work = 0;  // work buffer of the last advance into work (the forward's: W)
    int64_t j = 0;                    // reversals issued
    auto wait_for = [&](int k) -> int {
        if (k >= 0 && issued[k]) ADJ_CHECK(a, hipStreamWaitEvent(s2, ev[2 + k], 0));
        return SWE_OK;
    };
    // stream2 starts after the main stream's work so far (the forward's checkpoints)
    ADJ_CHECK(a, hipEventRecord(ev[4], c->stream));
    ADJ_CHECK(a, hipStreamWaitEvent(s2, ev[4], 0));
    for (size_t i = a->first_adj; i < a->acts.size(); ++i) {
        const Action &act = a->acts[i];
        const int k = (int)(j % 2);  // the pair of the next reversal
        int rc = SWE_OK;
        if (act.kind == kAdvance) {
            const int w = 1 - last_work;  // not the buffer the last reversal from work read
            rc = wait_for(work_reader[w]);
            if (!rc) rc = wait_for(k);  // the intermediate stage states use pair k
            if (!rc && act.dst >= 0) rc = wait_for(slot_reader[act.dst]);
            if (rc) return rc;
            work_reader[w] = -1;
            if (act.dst >= 0) slot_reader[act.dst] = -1;
            const State src = act.src < 0 ? work[last_work] : state_of(a, act.src);
            for (int64_t p = act.a; p < act.b && !rc; ++p) {
                const State in = p == act.a ? src : work[w];
                const State out =
                    (p == act.b - 1 && act.dst >= 0) ? state_of(a, act.dst) : work[w];
                rc = stages_12(a, in, pair[k][0], pair[k][1], p, s2, rows);
                if (!rc)
                    ADJ_CHECK(a, swe::launch_stage(3, c->bc, false, false, c->G,
                                                   fwd_args(a, pair[k][1], in, out, a->dts + p,
                                                            rows),
                                                   s2, nullptr));
            }
            if (rc) return rc;
            if (act.dst < 0) last_work = w;
            continue;
        }
        // reversal of step act.a from its start state
        const State base = act.src < 0 ? work[last_work] : state_of(a, act.src);
        rc = wait_for(k);
        if (!rc) rc = stages_12(a, base, pair[k][0], pair[k][1], act.a, s2, rows);
        if (rc) return rc;
        ADJ_CHECK(a, hipEventRecord(ev[k], s2));
        ADJ_CHECK(a, hipStreamWaitEvent(c->stream, ev[k], 0));
        rc = reverse_with(a, base, pair[k][0], pair[k][1], act.a);
        if (rc) return rc;
        ADJ_CHECK(a, hipEventRecord(ev[2 + k], c->stream));
        issued[k] = true;
        if (act.src < 0)
            work_reader[last_work] = k;
        else
            slot_reader[act.src] = k;
        ++j;
    }
    ADJ_CHECK(a, hipEventRecord(ev[4], s2));
    ADJ_CHECK(a, hipStreamWaitEvent(c->stream, ev[4], 0));
    // the events are released on return: drain both streams' use of them first
    ADJ_CHECK(a, hipStreamSynchronize(c->stream));
    return SWE_OK;
}

// The forward's run status words (swe::kStatusWords): a non-finite stage stopped it there.
int check_nonfinite(swe_hip_adj *a) {
    int w[swe::kStatusWords] = {};
    ADJ_CHECK(a, hipMemcpyAsync(w, a->ctx->flag, sizeof(w), hipMemcpyDeviceToHost,
                                a->ctx->stream));
    ADJ_CHECK(a, hipStreamSynchronize(a->ctx->stream));
    for (int s = 1; s <= 3; ++s)
        if (w[s] & swe::kStatusBad)
            return afail(a, SWE_ERR_NONFINITE,
                         "non-finite state in the forward pass at stage " + std::to_string(s) +
                             " of step " + std::to_string(w[0]));
    return SWE_OK;
}

int backward_common(swe_hip_adj *a, const double *lh, const double *lhu, const double *lhv,
                    double *gh, double *ghu, double *ghv, double *gb, hipMemcpyKind in_kind,
                    hipMemcpyKind out_kind) {
    if (!a || !lh || !lhu || !lhv || !gh || !ghu || !ghv) return SWE_ERR_INVALID;
    if (a->state != 1)
        return afail(a, SWE_ERR_INVALID, "backward needs a forward pass with fresh checkpoints");
    swe_hip_ctx *c = a->ctx;
    ADJ_CHECK(a, hipSetDevice(c->device));

    const auto t0 = std::chrono::steady_clock::now();
    const size_t bytes = a->n_total * sizeof(double);
    const double *lin[3] = {lh, lhu, lhv};
    for (int k = 0; k < 3; ++k) {
        ADJ_CHECK(a, hipMemcpyAsync(a->lam[0][k], lin[k], bytes, in_kind, c->stream));
        ADJ_CHECK(a, swe::launch_zero_ghosts(c->G, a->lam[0][k], c->stream));
    }
    ADJ_CHECK(a, swe::launch_fill_double(a->lam_b, 0.0, (long)a->n_total, c->stream));
    // the recomputing stage launches honour the first-bad-stage stop: clear what a
    // stepping call on the context may have left since the forward
    ADJ_CHECK(a, swe::launch_fill_int(c->flag, 0, swe::kStatusWords, c->stream));
    a->parity = 0;
    ADJ_CHECK(a, swe::launch_floor_mask(c->G, c->U[0], c->U[1], c->U[2], a->mask[0], c->stream));
    a->state = 2;
    if (a->records) {
        const int rc = backward_records(a);
        if (rc) return rc;
    } else if (a->stream2 && a->ovl[0]) {
        const int rc = backward_binomial_overlap(a);
        if (rc) return rc;
    } else {
        for (size_t i = a->first_adj; i < a->acts.size(); ++i) {
            const Action &act = a->acts[i];
            const int rc =
                act.kind == kAdvance ? advance(a, act, 0.0) : reverse_step(a, act.src, act.a);
            if (rc) return rc;
        }
    }
    double *gout[3] = {gh, ghu, ghv};
    for (int k = 0; k < 3; ++k)
        ADJ_CHECK(a, hipMemcpyAsync(gout[k], a->lam[0][k], bytes, out_kind, c->stream));
    if (gb) ADJ_CHECK(a, hipMemcpyAsync(gb, a->lam_b, bytes, out_kind, c->stream));
    ADJ_CHECK(a, hipStreamSynchronize(c->stream));
    a->bwd_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count();
    return check_nonfinite(a);
}

}  // namespace

extern "C" {

int swe_hip_adj_schedule(int64_t nsteps, int32_t nsnap, int64_t *advances, int32_t *max_reps,
                         int64_t *actions) {
    if (nsteps < 1 || nsnap < 1) return SWE_ERR_INVALID;
    const std::vector<Action> acts = make_schedule(nsteps, nsnap);
    std::vector<int32_t> reps((size_t)nsteps, 0);
    int64_t adv = 0;
    for (const Action &a : acts)
        if (a.kind == kAdvance) {
            adv += a.b - a.a;
            for (int64_t p = a.a; p < a.b; ++p) ++reps[(size_t)p];
        }
    if (advances) *advances = adv;
    if (max_reps) *max_reps = *std::max_element(reps.begin(), reps.end());
    if (actions) *actions = (int64_t)acts.size();
    return SWE_OK;
}

}  // extern "C"

namespace {

// Allocation shared by the create entry points: work buffers, adjoints, masks, dts.
int adj_alloc_common(swe_hip_adj *a) {
    const size_t bytes = a->n_total * sizeof(double);
    for (int i = 0; i < 3; ++i)
        if (hipMalloc(&a->W[i], bytes) != hipSuccess) return SWE_ERR_OOM;
    for (auto &l : a->lam)
        for (auto &p : l)
            if (hipMalloc(&p, bytes) != hipSuccess) return SWE_ERR_OOM;
    if (hipMalloc(&a->lam_b, bytes) != hipSuccess) return SWE_ERR_OOM;
    for (auto &m : a->mask)
        if (hipMalloc(&m, a->n_total) != hipSuccess) return SWE_ERR_OOM;
    if (hipMalloc(&a->dts, sizeof(double) * (size_t)a->nsteps) != hipSuccess ||
        hipMalloc(&a->partial, sizeof(double) * 1024) != hipSuccess ||
        hipMalloc(&a->counter, sizeof(long long)) != hipSuccess)
        return SWE_ERR_OOM;
    return SWE_OK;
}

// Binomial checkpoints: nsnap states of 3 arrays and the schedule.
int adj_setup_binomial(swe_hip_adj *a, int32_t nsnap) {
    a->records = false;
    a->nsnap = nsnap;
    a->per_slot = 3;
    if (hipMalloc(&a->slots, 3 * a->n_total * sizeof(double) * (size_t)nsnap) != hipSuccess)
        return SWE_ERR_OOM;
    a->acts = make_schedule(a->nsteps, nsnap);
    for (size_t i = 0; i < a->acts.size(); ++i)
        if (a->acts[i].kind == kAdjoint) {
            a->first_adj = i;
            break;
        }
    // the overlapped backward's working memory (9 arrays, like the work state and the
    // context's stage states not counted as stored states); without it the backward runs
    // the schedule on one stream
    if (a->nsteps >= 2) {
        bool ok = hipStreamCreateWithFlags(&a->stream2, hipStreamNonBlocking) == hipSuccess;
        for (auto &p : a->ovl)
            ok = ok && hipMalloc(&p, a->n_total * sizeof(double)) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            for (auto &p : a->ovl) {
                (void)hipFree(p);
                p = nullptr;
            }
            if (a->stream2) (void)hipStreamDestroy(a->stream2);
            a->stream2 = nullptr;
        }
    }
    return SWE_OK;
}

// Records plan (K >= 1).
int adj_setup_records(swe_hip_adj *a, const RecordPlan &pl) {
    a->records = true;
    a->rec_steps = pl.K;
    a->ckpt_every = pl.m;
    a->nck = pl.nck;
    a->rec_slots = pl.slots();
    a->nsnap = (int32_t)std::min<int64_t>(INT32_MAX, pl.K + pl.nck);
    a->per_slot = 3;
    const size_t arr = a->n_total * sizeof(double);
    if (hipMalloc(&a->rec, 9 * arr * (size_t)pl.slots()) != hipSuccess) return SWE_ERR_OOM;
    if (pl.nck >= 2 && hipStreamCreateWithFlags(&a->stream2, hipStreamNonBlocking) != hipSuccess)
        return SWE_ERR_HIP;
    if (pl.nck > 0 && hipMalloc(&a->slots, 3 * arr * (size_t)pl.nck) != hipSuccess)
        return SWE_ERR_OOM;
    return SWE_OK;
}

// The storage plan for `budget` bytes of stored states: the records plan with the fewest
// recomputed steps that fits, else binomial checkpoints in as many states as fit.
int adj_setup_budget(swe_hip_adj *a, double budget) {
    const double arr = (double)a->n_total * sizeof(double);
    const int64_t arrays = (int64_t)std::min(budget / arr, 9.0 * (double)a->nsteps + 9.0);
    const RecordPlan pl = plan_records(a->nsteps, arrays);
    if (pl.K >= 1) return adj_setup_records(a, pl);
    const int64_t fit = arrays / 3;
    if (fit < 1) return SWE_ERR_OOM;
    return adj_setup_binomial(a, (int32_t)std::min<int64_t>({a->nsteps, fit, INT32_MAX}));
}

int adj_begin(swe_hip_ctx *ctx, int64_t nsteps, swe_hip_adj **out, swe_hip_adj **made) {
    if (!ctx || !out || nsteps < 1) return SWE_ERR_INVALID;
    *out = nullptr;
    if (ctx->G.ng != 2 || ctx->scheme != SWE_SCHEME_OPTIMIZED) return SWE_ERR_INVALID;
    if (hipSetDevice(ctx->device) != hipSuccess) return SWE_ERR_HIP;
    auto *a = new swe_hip_adj();
    a->ctx = ctx;
    a->nsteps = nsteps;
    a->n_total = ctx->n_total;
    *made = a;
    return adj_alloc_common(a);
}

}  // namespace

extern "C" {

int swe_hip_adj_create(swe_hip_ctx *ctx, int64_t nsteps, int32_t nsnap, swe_hip_adj **out) {
    if (nsnap < -1) return SWE_ERR_INVALID;
    swe_hip_adj *a = nullptr;
    int rc = adj_begin(ctx, nsteps, out, &a);
    if (rc == SWE_OK) {
        if (nsnap == 0) {
            // automatic: the records plan (or binomial checkpoints) within 60% of the free
            // device memory
            size_t free_b = 0, total_b = 0;
            rc = hipMemGetInfo(&free_b, &total_b) == hipSuccess
                     ? adj_setup_budget(a, 0.6 * (double)free_b)
                     : SWE_ERR_HIP;
        } else if (nsnap == -1) {
            RecordPlan all;
            all.K = nsteps;
            rc = adj_setup_records(a, all);
        } else {
            rc = adj_setup_binomial(a, (int32_t)std::min<int64_t>(nsnap, nsteps));
        }
    }
    if (rc != SWE_OK) {
        swe_hip_adj_destroy(a);
        return rc;
    }
    *out = a;
    return SWE_OK;
}

int swe_hip_adj_create_budget(swe_hip_ctx *ctx, int64_t nsteps, double max_bytes,
                              swe_hip_adj **out) {
    if (!(max_bytes > 0.0)) return SWE_ERR_INVALID;
    swe_hip_adj *a = nullptr;
    int rc = adj_begin(ctx, nsteps, out, &a);
    if (rc == SWE_OK) rc = adj_setup_budget(a, max_bytes);
    if (rc != SWE_OK) {
        swe_hip_adj_destroy(a);
        return rc;
    }
    *out = a;
    return SWE_OK;
}

int swe_hip_adj_plan_budget(int64_t nsteps, int64_t n_total, double max_bytes,
                            int64_t *record_steps, int64_t *ckpt_every, int64_t *stored_states) {
    if (nsteps < 1 || n_total < 1 || !(max_bytes > 0.0)) return SWE_ERR_INVALID;
    const double arr = (double)n_total * sizeof(double);
    const int64_t arrays = (int64_t)std::min(max_bytes / arr, 9.0 * (double)nsteps + 9.0);
    const RecordPlan pl = plan_records(nsteps, arrays);
    int64_t K = pl.K, m = pl.m, states = pl.K + pl.nck;
    if (K < 1) {  // binomial fallback
        m = 0;
        states = std::min<int64_t>({nsteps, arrays / 3, INT32_MAX});
        if (states < 1) return SWE_ERR_OOM;
    }
    if (record_steps) *record_steps = K;
    if (ckpt_every) *ckpt_every = m;
    if (stored_states) *stored_states = states;
    return SWE_OK;
}

int swe_hip_adj_plan(const swe_hip_adj *a, int64_t *record_steps, int64_t *ckpt_every,
                     int64_t *stored_states, double *store_bytes, double *recomputed_steps) {
    if (!a) return SWE_ERR_INVALID;
    const double arr = (double)a->n_total * sizeof(double);
    int64_t K = 0, m = 0, states = a->nsnap;
    double bytes = 0.0, recomputed = 0.0;
    if (a->records) {
        K = a->rec_steps;
        m = a->ckpt_every;
        states = K + a->nck;
        RecordPlan pl;
        pl.K = K;
        pl.m = m;
        pl.nck = a->nck;
        bytes = arr * (double)pl.arrays();
        // each step before the tail once, the last step of every segment without stage 3
        recomputed = (double)(a->nsteps - K) - (double)a->nck / 3.0;
    } else {
        int64_t adv = 0;
        for (const Action &x : a->acts)
            if (x.kind == kAdvance) adv += x.b - x.a;
        bytes = arr * 3.0 * (double)a->nsnap;
        // advances beyond the forward sweep, plus U1 and U2 of every reversed step
        recomputed = (double)(adv - (a->nsteps - 1)) + 2.0 / 3.0 * (double)a->nsteps;
    }
    if (record_steps) *record_steps = K;
    if (ckpt_every) *ckpt_every = m;
    if (stored_states) *stored_states = states;
    if (store_bytes) *store_bytes = bytes;
    if (recomputed_steps) *recomputed_steps = recomputed;
    return SWE_OK;
}

void swe_hip_adj_destroy(swe_hip_adj *a) {
    if (!a) return;
    (void)hipSetDevice(a->ctx->device);
    (void)hipStreamSynchronize(a->ctx->stream);
    (void)hipFree(a->slots);
    (void)hipFree(a->rec);
    if (a->stream2) {
        (void)hipStreamSynchronize(a->stream2);
        (void)hipStreamDestroy(a->stream2);
    }
    for (auto *p : a->W) (void)hipFree(p);
    for (auto *p : a->ovl) (void)hipFree(p);
    for (auto &l : a->lam)
        for (auto *p : l) (void)hipFree(p);
    (void)hipFree(a->lam_b);
    for (auto *m : a->mask) (void)hipFree(m);
    (void)hipFree(a->dts);
    (void)hipFree(a->partial);
    (void)hipFree(a->counter);
    (void)hipFree(a->xf);
    (void)hipFree(a->xsel);
    delete a;
}

const char *swe_hip_adj_last_error(const swe_hip_adj *a) { return a ? a->err.c_str() : ""; }

int swe_hip_adj_forward(swe_hip_adj *a, int32_t dt_mode, double value, double *dt_hist) {
    if (!a) return SWE_ERR_INVALID;
    if (dt_mode == SWE_DT_GIVEN ? dt_hist == nullptr
                                : !(value > 0.0) || (dt_mode != SWE_DT_FIXED &&
                                                     dt_mode != SWE_DT_ADAPTIVE))
        return SWE_ERR_INVALID;
    swe_hip_ctx *c = a->ctx;
    ADJ_CHECK(a, hipSetDevice(c->device));
    const auto t0 = std::chrono::steady_clock::now();
    const size_t bytes = a->n_total * sizeof(double);
    double cfl = 0.0;
    if (dt_mode == SWE_DT_FIXED) {
        ADJ_CHECK(a, swe::launch_fill_double(a->dts, value, (long)a->nsteps, c->stream));
    } else if (dt_mode == SWE_DT_GIVEN) {
        for (int64_t p = 0; p < a->nsteps; ++p)
            if (!(dt_hist[p] > 0.0)) return afail(a, SWE_ERR_INVALID, "dt must be positive");
        ADJ_CHECK(a, hipMemcpyAsync(a->dts, dt_hist, sizeof(double) * (size_t)a->nsteps,
                                    hipMemcpyHostToDevice, c->stream));
    } else {
        cfl = value;
        ADJ_CHECK(a, swe::launch_fill_int(reinterpret_cast<int *>(a->counter), 0, 2, c->stream));
    }
    ADJ_CHECK(a, swe::launch_fill_int(c->flag, 0, swe::kStatusWords, c->stream));
    int rc = SWE_OK;
    if (a->records) {
        rc = forward_records(a, cfl);
    } else {
        // U^0 -> slot 0
        for (int k = 0; k < 3; ++k)
            ADJ_CHECK(a, hipMemcpyAsync(slot_ptr(a, 0, k), c->U[k], bytes,
                                        hipMemcpyDeviceToDevice, c->stream));
        // the schedule's first sweep: every action before the first reversal is an advance
        for (size_t i = 0; i < a->first_adj && rc == SWE_OK; ++i) rc = advance(a, a->acts[i], cfl);
        // the last step, from the start state the first reversal uses, into the context
        if (rc == SWE_OK) {
            const Action &first = a->acts[a->first_adj];
            const State cu{{c->U[0], c->U[1], c->U[2]}};
            rc = forward_step(a, state_of(a, first.src), cu, a->nsteps - 1, cfl);
        }
    }
    if (rc) return rc;
    if (dt_hist && dt_mode != SWE_DT_GIVEN)
        ADJ_CHECK(a, hipMemcpyAsync(dt_hist, a->dts, sizeof(double) * (size_t)a->nsteps,
                                    hipMemcpyDeviceToHost, c->stream));
    ADJ_CHECK(a, hipStreamSynchronize(c->stream));
    a->fwd_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count();
    rc = check_nonfinite(a);
    if (rc) return rc;
    a->state = 1;
    return SWE_OK;
}

int swe_hip_adj_backward(swe_hip_adj *a, const double *lh, const double *lhu, const double *lhv,
                         double *gh, double *ghu, double *ghv, double *gb) {
    return backward_common(a, lh, lhu, lhv, gh, ghu, ghv, gb, hipMemcpyHostToDevice,
                           hipMemcpyDeviceToHost);
}

int swe_hip_adj_backward_device(swe_hip_adj *a, const double *lh, const double *lhu,
                                const double *lhv, double *gh, double *ghu, double *ghv,
                                double *gb) {
    return backward_common(a, lh, lhu, lhv, gh, ghu, ghv, gb, hipMemcpyDeviceToDevice,
                           hipMemcpyDeviceToDevice);
}

int swe_hip_adj_set_launch(swe_hip_adj *a, int block_waves, int rows, int y_rows) {
    if (!a || rows < 0 || y_rows < 0) return SWE_ERR_INVALID;
    if (block_waves == 0) {
        if (!a->xf) {  // the x-face outputs of the two-kernel path, allocated on first use
            swe_hip_ctx *c = a->ctx;
            if (hipSetDevice(c->device) != hipSuccess) return afail(a, SWE_ERR_HIP, "hipSetDevice");
            const size_t xfaces = (size_t)(c->G.nx + 1) * c->G.stride;
            if (hipMalloc(&a->xf, 8 * xfaces * sizeof(double)) != hipSuccess ||
                hipMalloc(&a->xsel, xfaces + c->G.stride) != hipSuccess) {
                (void)hipFree(a->xf);
                a->xf = nullptr;
                a->xsel = nullptr;
                return afail(a, SWE_ERR_OOM, "x-face buffers of the two-kernel path");
            }
        }
        a->block_waves = 0;
        if (rows) a->tile_rows = rows;
        if (y_rows) a->y_rows = y_rows;
        return SWE_OK;
    }
    const int r = rows ? rows : a->fused_rows;
    if (!swe::adj_launch_ok(block_waves, r))
        return afail(a, SWE_ERR_INVALID,
                     "block_waves must be 2, 4 or 8 and its LDS tile within 160 KiB");
    a->block_waves = block_waves;
    a->fused_rows = r;
    return SWE_OK;
}

int swe_hip_adj_info(const swe_hip_adj *a, int32_t *nsnap, int32_t *records, double *forward_ms,
                     double *backward_ms) {
    if (!a) return SWE_ERR_INVALID;
    if (nsnap) *nsnap = a->nsnap;
    if (records) *records = a->records ? 1 : 0;
    if (forward_ms) *forward_ms = a->fwd_ms;
    if (backward_ms) *backward_ms = a->bwd_ms;
    return SWE_OK;
}

}  // extern "C"
