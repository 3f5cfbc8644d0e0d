// api.hip -- C ABI entry points: context, validation, binning and reordering.
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <new>

#include "lj_internal.h"

namespace lj {

static thread_local std::string g_create_err;

std::atomic<long long> g_alloc_budget{-1};
thread_local Arena* tl_arena = nullptr;

// Every device buffer a context owns (the workspace).
template <class F>
void for_each_buffer(lj_context* ctx, F&& f) {
    BinWork& b = ctx->bin;
    for (DevBuf* d : {&b.keys[0], &b.keys[1], &b.hist, &b.scan_tmp, &b.gaps, &b.offsets, &b.pos, &b.pos_alt,
                      &b.partials, &b.counters, &b.rows, &b.ccounts, &b.coffsets, &b.cand, &b.stage})
        f(*d);
    LbvhWork& l = ctx->lbvh;
    for (DevBuf* d : {&l.node_box, &l.node_child, &l.range, &l.chunks, &l.rope, &l.flags, &l.bucket_root}) f(*d);
    BatchWork& t = ctx->batch;
    for (DevBuf* d : {&t.sys, &t.hints, &t.npair, &t.ev, &t.wrec, &t.cand}) f(*d);
    f(ctx->tune_out);
}


int fail(lj_context* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    else g_create_err = buf;
    return code;
}

static float tri_lo() {
    float t = 1e-4f;
    if ((double)t > 1e-4) t = std::nextafter(t, 0.0f);
    return t;
}

// cudaLJ.cu:525-531: GridDim = (int)(BoxWidth / cutoff) in FP32.
static int grid_of(lj_context* ctx, const lj_params* p, GridGeom* g) {
    const float cw = p->cell_edge > 0.f ? p->cell_edge : p->rc;
    if (!(cw > 0.f) || !std::isfinite(cw)) return fail(ctx, LJ_ERR_INVALID_ARG, "cell edge must be finite and > 0");
    uint64_t nc = 1;
    for (int d = 0; d < 3; ++d) {
        if (!std::isfinite(p->box_lo[d]) || !std::isfinite(p->box_hi[d]))
            return fail(ctx, LJ_ERR_INVALID_ARG, "box bounds must be finite");
        const float L = p->box_hi[d] - p->box_lo[d];
        if (!(L > 0.f) || !std::isfinite(L)) return fail(ctx, LJ_ERR_INVALID_ARG, "box_hi must exceed box_lo on every axis");
        const float q = L / cw;
        if (!(q < 16777216.0f)) return fail(ctx, LJ_ERR_OVERFLOW, "more than 2^24 cells along one axis");
        int32_t gd = (int32_t)q;
        if (gd < 1) gd = 1;
        g->g[d] = gd;
        g->lo[d] = p->box_lo[d];
        g->gmax[d] = (float)(gd - 1);
        nc *= (uint64_t)gd;
    }
    if (nc > 0x7fffffffull) return fail(ctx, LJ_ERR_OVERFLOW, "cell count %llu exceeds 2^31-1", (unsigned long long)nc);
    g->cw[0] = g->cw[1] = g->cw[2] = cw;
    g->ncells = (uint32_t)nc;
    return LJ_OK;
}

int make_geom(lj_context* ctx, const lj_params* p, int64_t n, GridGeom* g, PairConst* c) {
    // every byte defined, padding included: the batched path compares plans bytewise to
    // decide whether to upload them again
    std::memset(g, 0, sizeof *g);
    std::memset(c, 0, sizeof *c);
    if (!p) return fail(ctx, LJ_ERR_INVALID_ARG, "params is NULL");
    if (n < 0 || n > 0x7fffffffll) return fail(ctx, LJ_ERR_INVALID_ARG, "particle count %lld out of range [0, 2^31-1]", (long long)n);
    if (!(p->rc > 0.f) || !std::isfinite(p->rc)) return fail(ctx, LJ_ERR_INVALID_ARG, "rc must be finite and > 0");
    if (!(p->sigma > 0.f) || !std::isfinite(p->sigma)) return fail(ctx, LJ_ERR_INVALID_ARG, "sigma must be finite and > 0");
    if (!std::isfinite(p->epsilon)) return fail(ctx, LJ_ERR_INVALID_ARG, "epsilon must be finite");
    if (!(p->softening2 >= 0.f) || !std::isfinite(p->softening2))
        return fail(ctx, LJ_ERR_INVALID_ARG, "softening2 must be finite and >= 0");
    if (p->predicate < LJ_PRED_AABB || p->predicate > LJ_PRED_TRIANGLE)
        return fail(ctx, LJ_ERR_INVALID_ARG, "unknown predicate %d", p->predicate);
    const float cw = p->cell_edge > 0.f ? p->cell_edge : p->rc;
    if (p->predicate != LJ_PRED_CUDALJ && cw < p->rc)
        return fail(ctx, LJ_ERR_INVALID_ARG, "cell_edge %g < rc %g would miss neighbours", (double)cw, (double)p->rc);
    LJ_TRY(grid_of(ctx, p, g));
    c->rc = p->rc;
    c->rc2 = p->rc * p->rc;
    c->sig2 = p->sigma * p->sigma;
    c->soft = p->predicate == LJ_PRED_CUDALJ ? p->softening2 : 0.0f;
    c->sig2d = (double)p->sigma * (double)p->sigma;
    c->softd = (double)c->soft;
    if (p->precision < LJ_PREC_MIXED || p->precision > LJ_PREC_FP32_COMP)
        return fail(ctx, LJ_ERR_INVALID_ARG, "unknown precision %d", p->precision);
    if (p->refine != 0 && p->refine != 1 && p->refine != 2 && p->refine != 4)
        return fail(ctx, LJ_ERR_INVALID_ARG, "refine must be 0, 1, 2 or 4 (got %d)", (int)p->refine);
    if (p->refine_x != 0 && p->refine_x != 1 && p->refine_x != 2 && p->refine_x != 4 && p->refine_x != 8 &&
        p->refine_x != 16)
        return fail(ctx, LJ_ERR_INVALID_ARG, "refine_x must be 0, 1, 2, 4, 8 or 16 (got %d)", (int)p->refine_x);
    if (p->refine_x != 0 && p->refine != 0 && p->refine_x < p->refine)
        return fail(ctx, LJ_ERR_INVALID_ARG, "refine_x (%d) must not be below refine (%d)", (int)p->refine_x,
                    (int)p->refine);
    if (p->block != 0 && p->block != 64 && p->block != 128 && p->block != 256)
        return fail(ctx, LJ_ERR_INVALID_ARG, "block must be 0, 64, 128 or 256 (got %d)", (int)p->block);
    if (p->depth != 0 && p->depth != 1 && p->depth != 8 && p->depth != 16 && p->depth != 32)
        return fail(ctx, LJ_ERR_INVALID_ARG, "depth must be 0, 1, 8, 16 or 32 (got %d)", (int)p->depth);
    c->tri_lo = tri_lo();
    // ray-model constants, rounded as the reference's host and raygen code rounds them
    // (SDK/optixMultiSphere/optixLJ.cu:95, optixSphere.cpp:157;
    // SDK/optixMultiTriangle/optixTriangle.cpp:184-205, optixLJ.cu:94-127)
    c->sph_h = p->rc * 0.81649658092f;
    c->sph_rad = p->rc * 0.81649658092f + 1e-5f;
    c->tri_hc = p->rc / 2;
    c->tri_ext = c->tri_hc + 0.00001f;
    c->tri_tmin = (double)1.19209290e-07F;
    c->tri_tmax = (double)(float)((double)(2 * c->tri_hc) + (double)(2 * c->tri_hc) * 0.00001);
    c->e4 = 4.0 * (double)p->epsilon;
    c->e24 = 24.0 * (double)p->epsilon;
    c->pred = p->predicate;
    return LJ_OK;
}

PhaseTimer::PhaseTimer(lj_context* ctx, int phase, hipStream_t s) : ctx_(ctx), phase_(phase), s_(s) {
    Profiler& P = ctx->prof;
    if (!P.on) return;
    auto& v = P.ev[phase];
    if (P.used[phase] == v.size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return;
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return;
        }
        v.emplace_back(a, b);
    }
    idx_ = (long)P.used[phase]++;
    (void)hipEventRecord(v[(size_t)idx_].first, s);
}

PhaseTimer::~PhaseTimer() {
    if (idx_ >= 0) (void)hipEventRecord(ctx_->prof.ev[phase_][(size_t)idx_].second, s_);
}

}  // namespace lj

using namespace lj;

extern "C" {

int lj_profile_enable(lj_context* ctx, int enable) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    ctx->prof.on = enable != 0;
    for (int q = 0; q < LJ_NUM_PHASES; ++q) ctx->prof.used[q] = 0;
    return LJ_OK;
}

int lj_profile_read(lj_context* ctx, double ms[3], unsigned long long calls[3]) {
    if (!ctx || !ms) return LJ_ERR_INVALID_ARG;
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    for (int q = 0; q < LJ_NUM_PHASES; ++q) {
        double sum = 0.0;
        for (size_t k = 0; k < ctx->prof.used[q]; ++k) {
            auto& e = ctx->prof.ev[q][k];
            LJ_HIP(ctx, hipEventSynchronize(e.second));
            float t = 0.f;
            LJ_HIP(ctx, hipEventElapsedTime(&t, e.first, e.second));
            sum += t;
        }
        ms[q] = sum;
        if (calls) calls[q] = ctx->prof.used[q];
    }
    return LJ_OK;
}

const char* lj_version(void) { return "lj_amd 0.1 (gfx950)"; }

static void release_streams(lj_context* ctx) {
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
    ctx->ev_join = ctx->ev_fork = nullptr;
    ctx->aux = nullptr;
}

int lj_context_create(lj_context** out, int device) {
    if (!out) return fail(nullptr, LJ_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return fail(nullptr, LJ_ERR_NO_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count)
        return fail(nullptr, LJ_ERR_INVALID_ARG, "device %d out of range [0,%d)", device, count);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return fail(nullptr, LJ_ERR_HIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    if (prop.warpSize != 64)
        return fail(nullptr, LJ_ERR_NO_DEVICE, "device %d has wavefront size %d; the kernels need 64", device, prop.warpSize);
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, LJ_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
    lj_context* ctx = new (std::nothrow) lj_context();
    if (!ctx) return fail(nullptr, LJ_ERR_OOM, "host allocation failed");
    ctx->device = device;
    ctx->num_cus = prop.multiProcessorCount;
    const int sr = scratch_setup(ctx, device);
    if (sr != LJ_OK) {
        fail(nullptr, sr, "%s", ctx->err.c_str());
        delete ctx;
        return sr;
    }
    // the second stream and its fork/join events (default priority; the caller's stream
    // may be the null stream, so the second one does not synchronise with it implicitly)
    {
        hipError_t se = hipSetDevice(device);
        if (se == hipSuccess) se = hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking);
        if (se == hipSuccess) se = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
        if (se == hipSuccess) se = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming);
        if (se != hipSuccess) {
            fail(nullptr, LJ_ERR_HIP, "second stream: %s", hipGetErrorString(se));
            release_streams(ctx);
            delete ctx;
            return LJ_ERR_HIP;
        }
    }
    // test hook: LJ_TEST_FAIL_ALLOC_AFTER=k lets k more device allocations succeed
    const char* inj = std::getenv("LJ_TEST_FAIL_ALLOC_AFTER");
    g_alloc_budget.store(inj && *inj ? std::strtoll(inj, nullptr, 10) : -1, std::memory_order_relaxed);
    *out = ctx;
    return LJ_OK;
}

static void release_all(lj_context* ctx) {
    for (auto& v : ctx->prof.ev)
        for (auto& e : v) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    for_each_buffer(ctx, [](DevBuf& d) { d.release(); });
}

int lj_context_destroy(lj_context* ctx) {
    if (!ctx) return LJ_OK;
    (void)hipSetDevice(ctx->device);
    release_all(ctx);
    release_streams(ctx);
    delete ctx;
    return LJ_OK;
}

int lj_workspace_generation(lj_context* ctx, uint64_t* generation) {
    if (!ctx || !generation) return LJ_ERR_INVALID_ARG;
    uint64_t g = 0;
    for_each_buffer(ctx, [&](DevBuf& d) { g += d.reallocs; });
    *generation = g;
    return LJ_OK;
}

const char* lj_last_error(const lj_context* ctx) {
    if (!ctx) return g_create_err.c_str();
    return ctx->err.c_str();
}

int lj_grid_dims(const lj_params* p, int32_t dims[3], int64_t* n_cells) {
    if (!p || !dims || !n_cells) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    const int r = make_geom(nullptr, p, 0, &g, &c);
    if (r) return r;
    for (int d = 0; d < 3; ++d) dims[d] = g.g[d];
    *n_cells = g.ncells;
    return LJ_OK;
}

int lj_bin(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
           uint32_t* cell_offsets, uint32_t* perm, unsigned long long* n_candidates, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    LJ_TRY(make_geom(ctx, p, n, &g, &c));
    if (n > 0 && (!x || !y || !z)) return fail(ctx, LJ_ERR_INVALID_ARG, "positions are NULL");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    LJ_TRY(bin_particles(ctx, x, y, z, n, g, s));
    if (cell_offsets)
        LJ_HIP(ctx, hipMemcpyAsync(cell_offsets, ctx->bin.offsets.ptr, ((size_t)g.ncells + 1) * 4,
                                   hipMemcpyDeviceToDevice, s));
    if (perm) LJ_TRY(extract_perm(ctx, n, perm, s));
    if (n_candidates) LJ_TRY(count_candidates(ctx, g, n_candidates, s));
    return LJ_OK;
}

}  // extern "C"
