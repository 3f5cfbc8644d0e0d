// tune.hip -- the autotuner: measured choice of the force-kernel shape.
//
// The shape of a force launch is the force-grid refinement (cells of edge
// cell_edge / K along y and z and cell_edge / Kx along x), the particles per workgroup
// and the depth of the per-lane LDS queues.  None of them changes which pairs are found; they trade
// candidate tests (fewer with finer cells) against cell-offset reads and loop
// overhead (more with finer cells), and occupancy against queue fill.  The best
// choice depends on the density and distribution of the particles, so it is
// measured, in two stages: every valid grid (K, Kx) with the default workgroup and no
// queue, then every workgroup size and queue depth on the fastest grid.  Each shape
// runs `reps` times on the caller's positions after one untimed call, timed with
// hipEvents on the caller's stream, and the fastest wins.
#include <vector>

#include "lj_stencil.h"

namespace lj {
namespace {

constexpr int kRefines[] = {1, 2, 4};
constexpr int kRefinesX[] = {1, 2, 4, 8, 16};
constexpr int kBlocks[] = {64, 128, 256};
constexpr int kDepths[] = {1, 8, 16, 32};  // 1: no queue

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Mean milliseconds of `reps` calls of fn after one untimed call; a negative value is
// the status of a failing call.
template <class Fn>
int time_calls(lj_context* ctx, EventPair& ev, hipStream_t s, int reps, Fn fn, float* ms) {
    LJ_TRY(fn());
    LJ_HIP(ctx, hipEventRecord(ev.a, s));
    for (int r = 0; r < reps; ++r) LJ_TRY(fn());
    LJ_HIP(ctx, hipEventRecord(ev.b, s));
    LJ_HIP(ctx, hipEventSynchronize(ev.b));
    float t = 0.f;
    LJ_HIP(ctx, hipEventElapsedTime(&t, ev.a, ev.b));
    *ms = t / (float)reps;
    return LJ_OK;
}

// Restores the profiler switch on every exit path: tuning calls are not profiled.
struct ProfilerOff {
    lj_context* ctx;
    bool was;
    explicit ProfilerOff(lj_context* c) : ctx(c), was(c->prof.on) { c->prof.on = false; }
    ~ProfilerOff() { ctx->prof.on = was; }
};

template <class Run>
int search(lj_context* ctx, hipStream_t s, int reps, bool cudalj, Run run, lj_tuning* best) {
    EventPair ev;
    LJ_HIP(ctx, hipEventCreate(&ev.a));
    LJ_HIP(ctx, hipEventCreate(&ev.b));
    ProfilerOff off(ctx);
    float ms = 0.f;
    LJ_TRY(time_calls(ctx, ev, s, reps, [&] { return run(0, 0, 0, 0); }, &ms));
    *best = lj_tuning{};
    best->ms_default = ms;
    best->ms_best = -1.f;
    auto consider = [&](int K, int Kx, int b, int d) {
        const int r = time_calls(ctx, ev, s, reps, [&] { return run(K, Kx, b, d); }, &ms);
        if (r != LJ_OK) return r;
        ++best->n_tried;
        if (best->ms_best < 0.f || ms < best->ms_best) {
            best->ms_best = ms;
            best->refine = K;
            best->refine_x = Kx;
            best->block = b;
            best->depth = d;
        }
        return LJ_OK;
    };
    // stage 1: grids (the CUDALJ mode keeps the reference's 27 cells)
    for (int K : kRefines)
        for (int Kx : kRefinesX) {
            if (Kx < K || (cudalj && (K != 1 || Kx != 1))) continue;
            const int r = consider(K, Kx, 256, 1);
            if (r == LJ_ERR_OVERFLOW) continue;  // grid too fine for this workload
            LJ_TRY(r);
        }
    if (best->n_tried == 0) return fail(ctx, LJ_ERR_INVALID_ARG, "no valid kernel shape for this workload");
    // stage 2: workgroup size and queue depth on the fastest grid
    const int K = best->refine, Kx = best->refine_x;
    for (int b : kBlocks)
        for (int d : kDepths) {
            if (b == 256 && d == 1) continue;  // timed in stage 1
            LJ_TRY(consider(K, Kx, b, d));
        }
    return LJ_OK;
}

}  // namespace
}  // namespace lj

using namespace lj;

extern "C" {

int lj_autotune(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, lj_params* p, int32_t reps,
                lj_tuning* out, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    LJ_TRY(make_geom(ctx, p, n, &g, &c));
    if (reps < 1 || reps > 1000) return fail(ctx, LJ_ERR_INVALID_ARG, "reps must be in [1, 1000] (got %d)", reps);
    if (n <= 0) return fail(ctx, LJ_ERR_INVALID_ARG, "autotuning needs particles (n = %lld)", (long long)n);
    if (!x || !y || !z) return fail(ctx, LJ_ERR_INVALID_ARG, "positions must not be NULL");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nn = (size_t)n;
    LJ_HIP(ctx, ctx->tune_out.grow(3 * nn * sizeof(float) + sizeof(lj_result) + 256));
    float* f = ctx->tune_out.as<float>();
    lj_result* res = reinterpret_cast<lj_result*>(ctx->tune_out.as<uint8_t>() + ((3 * nn * sizeof(float) + 255) & ~size_t(255)));
    lj_params q = *p;
    auto run = [&](int K, int Kx, int b, int d) {
        q.refine = (int16_t)K;
        q.refine_x = (int16_t)Kx;
        q.block = (int16_t)b;
        q.depth = (int16_t)d;
        GridGeom gq;
        PairConst cq;
        LJ_TRY(make_geom(ctx, &q, n, &gq, &cq));
        return compute_cellgrid(ctx, x, y, z, n, &q, gq, cq, f, f + nn, f + 2 * nn, nullptr, res, s);
    };
    lj_tuning best;
    LJ_TRY(search(ctx, s, reps, p->predicate == LJ_PRED_CUDALJ, run, &best));
    p->refine = (int16_t)best.refine;
    p->refine_x = (int16_t)best.refine_x;
    p->block = (int16_t)best.block;
    p->depth = (int16_t)best.depth;
    if (out) *out = best;
    return LJ_OK;
}

int lj_autotune_batched(lj_context* ctx, lj_system* sys, int32_t nsys, const float* x, const float* y, const float* z,
                        int32_t reps, lj_tuning* out, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    if (nsys <= 0 || !sys) return fail(ctx, LJ_ERR_INVALID_ARG, "autotuning needs a non-empty system array");
    if (reps < 1 || reps > 1000) return fail(ctx, LJ_ERR_INVALID_ARG, "reps must be in [1, 1000] (got %d)", reps);
    int64_t ntot = 0;
    for (int32_t i = 0; i < nsys; ++i) {
        if (sys[i].n < 0) return fail(ctx, LJ_ERR_INVALID_ARG, "system %d has n < 0", i);
        ntot += sys[i].n;
    }
    if (ntot <= 0 || ntot > 0x7fffffffll) return fail(ctx, LJ_ERR_INVALID_ARG, "batch particle count out of range");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nn = (size_t)ntot;
    const size_t fbytes = (3 * nn * sizeof(float) + 255) & ~size_t(255);
    LJ_HIP(ctx, ctx->tune_out.grow(fbytes + (size_t)nsys * sizeof(lj_result)));
    float* f = ctx->tune_out.as<float>();
    lj_result* res = reinterpret_cast<lj_result*>(ctx->tune_out.as<uint8_t>() + fbytes);
    std::vector<lj_system> q(sys, sys + nsys);
    auto run = [&](int K, int Kx, int b, int d) {
        for (auto& t : q) {
            t.params.refine = (int16_t)K;
            t.params.refine_x = (int16_t)Kx;
            t.params.block = (int16_t)b;
            t.params.depth = (int16_t)d;
        }
        return lj_compute_batched(ctx, q.data(), nsys, x, y, z, f, f + nn, f + 2 * nn, nullptr, res, stream);
    };
    lj_tuning best;
    LJ_TRY(search(ctx, s, reps, sys[0].params.predicate == LJ_PRED_CUDALJ, run, &best));
    for (int32_t i = 0; i < nsys; ++i) {
        sys[i].params.refine = (int16_t)best.refine;
        sys[i].params.refine_x = (int16_t)best.refine_x;
        sys[i].params.block = (int16_t)best.block;
        sys[i].params.depth = (int16_t)best.depth;
    }
    if (out) *out = best;
    return LJ_OK;
}

}  // extern "C"
