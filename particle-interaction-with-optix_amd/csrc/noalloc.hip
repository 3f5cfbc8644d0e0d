// noalloc.hip -- lj_compute_noalloc: lj_compute with its workspace in a fixed,
// file-scope device array instead of the context's grow-only buffers.
//
// For callers that cannot allocate device memory during a call (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The kernels and
// the pipeline are lj_compute's (compute_cellgrid); only where the workspace comes
// from differs: every buffer the pipeline grows is carved, for the duration of the
// call, from `scratch` by a bump allocator, so the call makes no hipMalloc / hipFree
// and no other allocating call (profiling is suspended for it: event creation
// allocates).  Sizes beyond the array's capacity are rejected up front.
//
// The array is one per GPU (each device holds its own copy, at its own address) and
// lives until the process ends.  Every call on a GPU shares it, and its contents are
// whatever the last call left: the pipeline clears what must start at zero itself, as
// it does for the context's buffers.  Calls that overlap on one GPU (two streams, two
// host threads, two contexts) race on the array and must be serialised by the caller.
#include <mutex>

#include "lj_stencil.h"

namespace lj {

constexpr size_t kScratchBytes = (size_t)LJ_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size is a multiple of the carve alignment");
// The outputs of the largest input fit in the L2s, so the pipeline stores them directly
// and never needs the staging array of larger calls.
static_assert(!(non_temporal_outputs((int64_t)LJ_NOALLOC_MAX_PARTICLES)), "no staged outputs under lj_compute_noalloc");

// The workspace of lj_compute_noalloc: 64 MiB per GPU, enough for the largest input
// it takes (2^20 particles on a force grid of 2^21 cells needs about 57 MiB).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Device address of this device's copy of `scratch`, looked up once per device with
// the device current.  The array's name in host code is only a host-side handle:
// passing it to a kernel would fault, so kernels get this address.
int scratch_setup(lj_context* ctx, int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(ctx, LJ_ERR_INVALID_ARG, "device %d beyond the %d the library tracks", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        LJ_HIP(ctx, hipGetDevice(&cur));
        if (cur != device) LJ_HIP(ctx, hipSetDevice(device));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr)
            return fail(ctx, LJ_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                        hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    ctx->scratch = g_scratch_ptr[device];
    return LJ_OK;
}

}  // namespace lj

using namespace lj;

extern "C" int lj_compute_noalloc(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
                                  const lj_params* p, float* fx, float* fy, float* fz, float* e, lj_result* result,
                                  void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    LJ_TRY(make_geom(ctx, p, n, &g, &c));
    if (n > 0 && (!x || !y || !z || !fx || !fy || !fz))
        return fail(ctx, LJ_ERR_INVALID_ARG, "positions and forces must not be NULL");
    if (!ctx->scratch) return fail(ctx, LJ_ERR_HIP, "the scratch array of device %d was not set up", ctx->device);
    if (n > LJ_NOALLOC_MAX_PARTICLES)
        return fail(ctx, LJ_ERR_CAPACITY, "lj_compute_noalloc takes at most %d particles (got %lld)",
                    LJ_NOALLOC_MAX_PARTICLES, (long long)n);
    // the force grid of the shape this call will use
    const Shape sh = choose_shape(p, n, g.ncells);
    GridGeom fg = g;
    if (sh.refine > 1 || sh.refine_x > 1) LJ_TRY(refine_geom(ctx, g, p, sh.refine_x, sh.refine, &fg));
    if (fg.ncells > (uint32_t)LJ_NOALLOC_MAX_CELLS)
        return fail(ctx, LJ_ERR_CAPACITY, "lj_compute_noalloc takes at most %d force-grid cells (got %u)",
                    LJ_NOALLOC_MAX_CELLS, fg.ncells);
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    // the pipeline runs on an empty workspace whose buffers are carved from scratch;
    // the context's own buffers are put back afterwards, untouched
    Arena arena{static_cast<unsigned char*>(ctx->scratch), kScratchBytes, 0};
    BinWork carved;
    std::swap(ctx->bin, carved);
    const bool prof = ctx->prof.on;
    ctx->prof.on = false;
    Arena* const prev = tl_arena;
    tl_arena = &arena;
    const int rc = compute_cellgrid(ctx, x, y, z, n, p, g, c, fx, fy, fz, e, result, (hipStream_t)stream);
    tl_arena = prev;
    ctx->prof.on = prof;
    std::swap(ctx->bin, carved);
    if (rc == LJ_ERR_OOM)
        return fail(ctx, LJ_ERR_CAPACITY, "lj_compute_noalloc: the workspace exceeds the %zu-byte scratch array",
                    kScratchBytes);
    return rc;
}
