// lj_internal.h -- context, workspace and launch helpers shared by the HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "lj_amd.h"
#include "lj_device.h"

struct lj_context;

namespace lj {

// Allocation-failure injection for tests: when >= 0, the number of device
// allocations that may still succeed in this process before every later one reports
// hipErrorOutOfMemory (without calling hipMalloc).  Set by lj_context_create from the
// environment variable LJ_TEST_FAIL_ALLOC_AFTER; -1 (the default) turns it off.
extern std::atomic<long long> g_alloc_budget;

inline bool alloc_fault_injected() {
    long long b = g_alloc_budget.load(std::memory_order_relaxed);
    while (b > 0 && !g_alloc_budget.compare_exchange_weak(b, b - 1, std::memory_order_relaxed)) {
    }
    return b == 0;
}

// A fixed block of device memory that grow() carves buffers from, front to back,
// instead of allocating (lj_compute_noalloc points tl_arena at it for one call).
struct Arena {
    unsigned char* base;
    size_t size, used;
};
extern thread_local Arena* tl_arena;

// A grow-only device buffer.  grow() reallocates only when the request exceeds the
// current capacity, so steady-state calls of one size allocate nothing.  `reallocs`
// counts the allocations, so a caller holding the buffer's address (a captured
// graph) can tell that it moved (lj_workspace_generation).
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    uint64_t reallocs = 0;
    hipError_t grow(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (Arena* a = tl_arena) {
            // carve from the arena; the carved buffers are dropped with the call
            const size_t cap = ((need < 256 ? 256 : need) + 255) / 256 * 256;
            if (cap > a->size - a->used) return hipErrorOutOfMemory;
            ptr = a->base + a->used;
            a->used += cap;
            bytes = cap;
            return hipSuccess;
        }
        if (alloc_fault_injected()) return hipErrorOutOfMemory;
        if (ptr) {
            hipError_t e = hipFree(ptr);
            if (e != hipSuccess) return e;
            ptr = nullptr;
            bytes = 0;
        }
        size_t cap = need < 256 ? 256 : need;
        ++reallocs;
        hipError_t e = hipMalloc(&ptr, cap);
        if (e != hipSuccess) { ptr = nullptr; return e; }
        bytes = cap;
        return hipSuccess;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Status carrier: the first failing check records its message on the context.
struct Status {
    int code = LJ_OK;
    explicit operator bool() const { return code != LJ_OK; }
};

// Workspace of one binned system (or of a concatenated batch).
// Entries past the last sorted particle in BinWork::pos (never written, read but unused).
constexpr int kPosSpare = 3;

struct BinWork {
    DevBuf keys[2];      // u32 radix keys (ping-pong)
    DevBuf hist;         // radix digit histograms
    DevBuf scan_tmp;     // partial sums of the device-wide scans
    DevBuf gaps;         // count + list of long empty-cell stretches for the cell offsets
    DevBuf offsets;      // u32 per-cell offsets (ncells+1)
    DevBuf pos;          // float4 sorted positions, w = original index bits
    DevBuf pos_alt;      // float4 ping-pong partner of pos during the sort
    DevBuf partials;     // per-block partial sums (double) for deterministic totals
    DevBuf counters;     // u64 counters
    DevBuf rows;         // u64 pair-list row offsets
    DevBuf ccounts;      // u32 counts of the cell_edge grid when the force grid is refined
    DevBuf coffsets;     // u32 offsets of the cell_edge grid (ncells+1)
    DevBuf cand;         // u64 per-workgroup partial candidate counts
    DevBuf stage;        // float4 per particle: outputs staged at the caller index
    int cur = 0;         // which keys buffer holds the sorted keys
    const uint32_t* sorted_keys() const { return keys[cur].as<uint32_t>(); }
};

// The LBVH keeps its leaves (positions in Morton order) in BinWork::pos and its
// Morton codes in BinWork::keys.
struct LbvhWork {
    DevBuf node_box;      // float4 x2 per internal node (lo, hi)
    DevBuf node_child;    // int2 per internal node
    DevBuf range;         // int2 per internal node: first and last leaf it covers
    DevBuf chunks;        // float4 x2 bounds of aligned leaf chunks (levels of 4, 16, ..., 2^24 leaves)
    DevBuf rope;          // int per leaf: the node after every subtree ending at that leaf
    DevBuf flags;         // the traversal overflow flag
    DevBuf bucket_root;   // uint8 per internal node: 1 when its parent scans it as a bucket
};

struct BatchWork {
    DevBuf sys;           // device copy of the plan: per-system descriptors and cell bases
    DevBuf hints;         // search hints of the plan's three base arrays (rebuilt with the plan)
    DevBuf npair;         // u32 accepted-pair count of each sorted slot in a wave shared by systems
    DevBuf ev;            // double2 energy and virial of each sorted slot in a wave shared by systems
    DevBuf wrec;          // one WaveTotals per wave of 64 sorted slots
    DevBuf cand;          // u64 candidate count per system, counted on the second stream
    std::vector<uint8_t> plan;  // host copy of the plan last uploaded to sys
};

}  // namespace lj

namespace lj {

// Event pairs per phase; reused across enable/read cycles.
struct Profiler {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[LJ_NUM_PHASES];
    size_t used[LJ_NUM_PHASES] = {0, 0, 0};
};

}  // namespace lj

struct lj_context {
    int device = 0;
    lj::Profiler prof;
    std::string err;
    int num_cus = 0;
    lj::BinWork bin;
    lj::LbvhWork lbvh;
    lj::BatchWork batch;
    lj::DevBuf tune_out;  // outputs of the autotuner's timed calls
    void* scratch = nullptr;  // device address of this device's lj_compute_noalloc array
    // A second stream (default priority) for work that can run beside the force pass
    // (the batched call's candidate counts), forked from and joined into the caller's
    // stream with these two events.
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace lj {

int fail(lj_context* ctx, int code, const char* fmt, ...);

// Looks up (once per device) the device address of lj_compute_noalloc's scratch array
// and stores it in ctx->scratch (noalloc.hip).
int scratch_setup(lj_context* ctx, int device);

#define LJ_HIP(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e__ = (expr);                                                            \
        if (e__ != hipSuccess)                                                              \
            return ::lj::fail((ctx), e__ == hipErrorOutOfMemory ? LJ_ERR_OOM : LJ_ERR_HIP,  \
                              "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e__),       \
                              __FILE__, __LINE__);                                          \
    } while (0)

#define LJ_TRY(expr)              \
    do {                          \
        int r__ = (expr);         \
        if (r__ != LJ_OK) return r__; \
    } while (0)

// Validates params and derives the grid and pair constants.  n is the particle
// count of the system (for the index-width checks).
int make_geom(lj_context* ctx, const lj_params* p, int64_t n, GridGeom* g, PairConst* c);

// Device-wide exclusive scan of n u32 values (in may alias out).  Writes the total
// to *total_dev (nullable).  Uses ctx->bin.scan_tmp.
int scan_u32(lj_context* ctx, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev,
             hipStream_t s);

// Exclusive scan of n u32 values into u64 out[0..n), with out[n] = total.
int scan_u32_to_u64(lj_context* ctx, const uint32_t* in, unsigned long long* out, int64_t n, hipStream_t s);

// Cell-grid pipelines (force.hip).
// Where a force launch also leaves the particles in cell order (lj_compute_sorted):
// positions and original indices by sorted slot, outputs by sorted slot too.  All null:
// outputs by original index.
struct SortedOut {
    float* x = nullptr;
    float* y = nullptr;
    float* z = nullptr;
    uint32_t* perm = nullptr;
};
int compute_cellgrid(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
                     const GridGeom& g, const PairConst& c, float* fx, float* fy, float* fz, float* e,
                     lj_result* result, hipStream_t s, const SortedOut& so = SortedOut{});

// Key functor of the cell grid (for sort_particles in lj_sort.h).
struct CellKeyOp {
    GridGeom g;
    __device__ uint32_t operator()(float x, float y, float z, uint32_t) const { return cell_key(g, x, y, z); }
};

// Cell counts and offsets (ncells+1) of ctx->bin from the sorted keys of a sort_particles.
int cell_offsets_from_sorted(lj_context* ctx, int64_t n, uint32_t ncells, hipStream_t s);

// perm[k] = original index of sorted slot k (from ctx->bin.pos after a sort).
int extract_perm(lj_context* ctx, int64_t n, uint32_t* perm, hipStream_t s);

// res->e_total / res->virial from per-block FP64 partials (pairs of doubles), fixed order.
// partials must hold room for nblocks + nblocks/1024 + 1 pairs (stage-1 results).
void finalize_energy(double* partials, unsigned nblocks, lj_result* res, hipStream_t s);

// Bins n particles: fills ctx->bin.offsets (ncells+1), ctx->bin.pos (sorted float4 with
// the original index in w) and the sorted cell keys (ctx->bin.sorted_keys()).
int bin_particles(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
                  const GridGeom& g, hipStream_t s);

// Candidate count of the 27-cell stencil from ctx->bin.offsets, into *out (device u64),
// which is overwritten (not accumulated).
int count_candidates(lj_context* ctx, const GridGeom& g, unsigned long long* out, hipStream_t s);
// Same from explicit offsets of grid g.
int count_candidates_from(lj_context* ctx, const GridGeom& g, const uint32_t* offsets, unsigned long long* out,
                          hipStream_t s);

void fill_u64(unsigned long long* p, uint32_t n, unsigned long long v, hipStream_t s);
// fx, fy, fz, e[i] = stage[i] for i < n (e nullable): outputs staged as one float4 per
// particle at its caller index, copied out with consecutive loads and stores.
void split_outputs(const float4* stage, uint32_t n, float* fx, float* fy, float* fz, float* e, hipStream_t s);
void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);

// Records a phase's start event on construction and its stop event on destruction
// when profiling is on.
class PhaseTimer {
  public:
    PhaseTimer(lj_context* ctx, int phase, hipStream_t s);
    ~PhaseTimer();

  private:
    lj_context* ctx_;
    int phase_;
    hipStream_t s_;
    long idx_ = -1;
};

inline unsigned div_up(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace lj
