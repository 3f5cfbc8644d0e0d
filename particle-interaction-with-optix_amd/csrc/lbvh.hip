// lbvh.hip -- software linear BVH over the particles, the MI355X counterpart of the
// reference's OptiX AABB variant (CDNA has no ray-tracing units).
//
// Reference: optixMultiSphereAABB builds one AABB [p - rc, p + rc] per particle
// (SDK/optixMultiSphereAABB/optixSphere.cpp:204-233), builds a BVH over them with
// optixAccelBuild (:235-289), and launches one zero-length ray per particle from its
// position (:511-520, optixLJ.cu:49-88); __intersection__sphere accepts a box when
// 1e-16 <= d2 <= c^2 and adds the LJ energy (optixLJ.cu:134-176).
//
// Here:
//  1. 30-bit Morton codes of the positions in the params box, stable radix sort
//     (ties keep index order, so the sorted (code, index) keys are unique);
//  2. the hierarchy of Karras, "Maximizing Parallelism in the Construction of BVHs,
//     Octrees, and k-d Trees", HPG 2012 (one thread per internal node);
//  3. bottom-up bounds: each leaf walks to the root, the second child to arrive at a
//     node (acquire/release counter) merges the two child boxes;
//  4. one thread per particle (in Morton order) traverses with a stack, testing the
//     query box [p - rc', p + rc'] against node point-bounds, then the canonical
//     predicate and the pair arithmetic of the cell grid on each leaf it reaches.
// A node stores the bounds of its particle positions; testing the query box of
// half-width rc' = rc (1 + 2^-18) against them is the point-in-AABB test of the
// reference with the rounding of p +- rc moved to the safe side: every pair the
// predicate can accept has |d_axis| <= rc (1 + 2^-20).
#include "lj_internal.h"
#include "lj_pair.h"
#include "lj_sort.h"

namespace lj {

namespace {

constexpr int kStack = 64;

struct Box {
    float lo[3], hi[3];
};

__device__ inline uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ inline uint32_t quant10(float x, float lo, float inv) {
    float q = (x - lo) * inv;
    q = fminf(fmaxf(q, 0.0f), 1023.0f);
    return (uint32_t)q;
}

// 30-bit Morton code of a position, quantised to 1024 steps per axis of the box.
struct MortonKeyOp {
    float lo[3], inv[3];
    __device__ uint32_t operator()(float x, float y, float z, uint32_t) const {
        return (spread10(quant10(x, lo[0], inv[0])) << 2) | (spread10(quant10(y, lo[1], inv[1])) << 1) |
               spread10(quant10(z, lo[2], inv[2]));
    }
};

// Common-prefix length of sorted keys i and j (Karras' delta); equal codes fall back
// to the indices, which makes every key distinct.
__device__ inline int delta(const uint32_t* __restrict__ code, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = code[i], b = code[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

// Node numbering: internal nodes 0..n-2 (0 is the root), leaf k is node n-1+k.
__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ code, int n, int2* __restrict__ child,
                                                int* __restrict__ parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(code, n, i, i + 1) - delta(code, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(code, n, i, i - d);
    int lmax = 2;
    while (delta(code, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(code, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(code, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(code, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int left = (lo == gamma) ? (n - 1 + gamma) : gamma;
    const int right = (hi == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    child[i] = make_int2(left, right);
    parent[left] = i;
    parent[right] = i;
}

__global__ __launch_bounds__(256) void k_fill_i32(int* __restrict__ p, uint32_t n, int v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Bottom-up bounds.  Node boxes: box[node] = (lo.xyz, hi.xyz) as two float4.
__global__ __launch_bounds__(256) void k_refit(const float4* __restrict__ leaf, int n, const int2* __restrict__ child,
                                               const int* __restrict__ parent, float4* __restrict__ box,
                                               int* __restrict__ arrived, int* __restrict__ overflow) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 p = leaf[k];
    int node = n - 1 + k;
    box[2 * node] = make_float4(p.x, p.y, p.z, 0.f);
    box[2 * node + 1] = make_float4(p.x, p.y, p.z, 0.f);
    int par = parent[node];
    for (int depth = 0; par >= 0; ++depth) {
        if (depth >= 2 * kStack) {  // deeper than any well-formed tree: stop and flag
            atomicOr(overflow, 2);
            return;
        }
        // release our box, acquire the sibling's: the second arrival proceeds
        const int prev = __hip_atomic_fetch_add(&arrived[par], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == 0) return;
        const int2 c = child[par];
        const float4 a0 = box[2 * c.x], a1 = box[2 * c.x + 1];
        const float4 b0 = box[2 * c.y], b1 = box[2 * c.y + 1];
        box[2 * par] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.f);
        box[2 * par + 1] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.f);
        node = par;
        par = parent[node];
    }
}

// Does the query box [p - h, p + h] meet node bounds [lo, hi]?  Differences in FP32
// are monotone in their operands, so the test never rejects a point within h.
__device__ inline bool overlaps(float4 p, float h, float4 lo, float4 hi) {
    return (lo.x - p.x) <= h && (p.x - hi.x) <= h && (lo.y - p.y) <= h && (p.y - hi.y) <= h && (lo.z - p.z) <= h &&
           (p.z - hi.z) <= h;
}

// Visits every leaf whose position lies within the query box of leaf k, calling
// f(slot, leaf position) for each one except k itself.  Returns false on stack overflow
// (the tree is deeper than kStack levels, impossible for 30-bit codes plus 32-bit
// indices, but checked).
template <class F>
__device__ inline bool traverse(const float4* __restrict__ leaf, const int2* __restrict__ child,
                                const float4* __restrict__ box, int n, int k, float4 p, float h, unsigned long long& tests,
                                F&& f) {
    if (n == 1) return true;
    int stack[kStack];
    int sp = 0;
    stack[sp++] = 0;
    int visits = 0;
    while (sp > 0) {
        if (++visits > n) return false;  // a well-formed tree has n - 1 internal nodes
        const int node = stack[--sp];
        const int2 c = child[node];
        const int kids[2] = {c.x, c.y};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int ch = kids[q];
            if (ch >= n - 1) {
                const int slot = ch - (n - 1);
                const float4 pj = leaf[slot];
                if ((pj.x - p.x) <= h && (p.x - pj.x) <= h && (pj.y - p.y) <= h && (p.y - pj.y) <= h &&
                    (pj.z - p.z) <= h && (p.z - pj.z) <= h) {
                    ++tests;
                    if (slot != k) f(slot, pj);
                }
            } else if (overlaps(p, h, box[2 * ch], box[2 * ch + 1])) {
                if (sp >= kStack) return false;
                stack[sp++] = ch;
            }
        }
    }
    return true;
}

template <int PRED, int PREC, bool VIR>
__global__ __launch_bounds__(256) void k_lbvh_force(const float4* __restrict__ leaf, const int2* __restrict__ child,
                                                    const float4* __restrict__ box, int n, float h, PairConst c,
                                                    float* __restrict__ fx, float* __restrict__ fy,
                                                    float* __restrict__ fz, float* __restrict__ e,
                                                    double* __restrict__ partials, lj_result* __restrict__ res,
                                                    int* __restrict__ overflow) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    PairSums<PREC, false, VIR> acc;
    unsigned long long np = 0, tests = 0;
    double esum = 0.0, vsum = 0.0;
    if (k < n) {
        const float4 pi = leaf[k];
        acc.set_i(pi);
        const bool ok = traverse(leaf, child, box, n, k, pi, h, tests, [&](int, float4 pj) {
            const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const float d2 = dist2(dx, dy, dz);
            if (!accept<PRED>(c, d2)) return;
            ++np;
            acc.add(c, pj, dx, dy, dz, d2);
        });
        if (!ok) atomicOr(overflow, 1);
        const uint32_t i = __float_as_uint(pi.w);
        fx[i] = (float)(-c.e24 * (double)acc.ax.value());
        fy[i] = (float)(-c.e24 * (double)acc.ay.value());
        fz[i] = (float)(-c.e24 * (double)acc.az.value());
        const double ei = c.e4 * (double)acc.ae.value();
        if (e) e[i] = (float)ei;
        esum = ei;
        vsum = c.e24 * (double)acc.av.value();
    }
    block_sum2(esum, vsum);
    const unsigned long long bp = block_sum_u64(np);
    const unsigned long long bt = block_sum_u64(tests);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = esum;
        partials[2 * blockIdx.x + 1] = vsum;
        if (bp) atomicAdd(&res->n_pairs, bp);
        if (bt) atomicAdd(&res->n_candidates, bt);
    }
}

template <int PRED>
__global__ __launch_bounds__(256) void k_lbvh_count(const float4* __restrict__ leaf, const int2* __restrict__ child,
                                                    const float4* __restrict__ box, int n, float h, PairConst c,
                                                    uint32_t* __restrict__ rowcnt, int* __restrict__ overflow) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 pi = leaf[k];
    uint32_t cnt = 0;
    unsigned long long tests = 0;
    const bool ok = traverse(leaf, child, box, n, k, pi, h, tests, [&](int, float4 pj) {
        if (accept<PRED>(c, dist2(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z))) ++cnt;
    });
    if (!ok) atomicOr(overflow, 1);
    rowcnt[__float_as_uint(pi.w)] = cnt;
}

template <int PRED>
__global__ __launch_bounds__(256) void k_lbvh_fill(const float4* __restrict__ leaf, const int2* __restrict__ child,
                                                   const float4* __restrict__ box, int n, float h, PairConst c,
                                                   const unsigned long long* __restrict__ rows,
                                                   uint32_t* __restrict__ cols, int* __restrict__ overflow) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 pi = leaf[k];
    const uint32_t i = __float_as_uint(pi.w);
    const unsigned long long r0 = rows[i], r1 = rows[i + 1];
    unsigned long long w = r0;
    unsigned long long tests = 0;
    const bool ok = traverse(leaf, child, box, n, k, pi, h, tests, [&](int, float4 pj) {
        if (w < r1 && accept<PRED>(c, dist2(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z))) cols[w++] = __float_as_uint(pj.w);
    });
    if (!ok) atomicOr(overflow, 1);
    for (unsigned long long a = r0 + 1; a < w; ++a) {
        const uint32_t v = cols[a];
        unsigned long long b = a;
        while (b > r0 && cols[b - 1] > v) {
            cols[b] = cols[b - 1];
            --b;
        }
        cols[b] = v;
    }
}

}  // namespace

// The overflow flag lives after the n arrival counters in ctx->lbvh.flags.
static int* overflow_flag(lj_context* ctx, int64_t n) { return ctx->lbvh.flags.as<int>() + n + 1; }

// Builds the tree over n particles into ctx->lbvh.  h is the query half-width.
static int lbvh_build(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
                      hipStream_t s) {
    LbvhWork& L = ctx->lbvh;
    BinWork& B = ctx->bin;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    LJ_HIP(ctx, L.node_child.grow(nn * sizeof(int2)));
    LJ_HIP(ctx, L.parent.grow(2 * nn * sizeof(int)));
    LJ_HIP(ctx, L.flags.grow(nn * sizeof(int) + 64));
    LJ_HIP(ctx, L.node_box.grow(2 * nn * 2 * sizeof(float4)));
    MortonKeyOp op;
    for (int d = 0; d < 3; ++d) {
        op.lo[d] = p->box_lo[d];
        op.inv[d] = 1024.0f / (p->box_hi[d] - p->box_lo[d]);
    }
    // the sorted payload is the leaf array: positions in Morton order, w = original index
    LJ_TRY(sort_particles(ctx, x, y, z, n, op, 30, s));
    if (n == 0) return LJ_OK;
    const unsigned nb = div_up(n, 256);
    const uint32_t* code = B.sorted_keys();
    int* parent = L.parent.as<int>();
    int* arrived = L.flags.as<int>();
    hipLaunchKernelGGL(k_fill_i32, dim3(div_up(2 * n, 256)), dim3(256), 0, s, parent, (uint32_t)(2 * n), -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(div_up(n + 1, 256)), dim3(256), 0, s, arrived, (uint32_t)n + 1, 0);
    if (n > 1)
        hipLaunchKernelGGL(k_karras, dim3(div_up(n - 1, 256)), dim3(256), 0, s, code, (int)n, L.node_child.as<int2>(),
                           parent);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, overflow_flag(ctx, n), 1u, 0);
    hipLaunchKernelGGL(k_refit, dim3(nb), dim3(256), 0, s, (const float4*)B.pos.as<float4>(), (int)n,
                       (const int2*)L.node_child.as<int2>(), (const int*)parent, L.node_box.as<float4>(), arrived,
                       overflow_flag(ctx, n));
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

static float query_half_width(const lj_params* p) { return p->rc * (1.0f + 0x1p-18f); }

static int check_lbvh_params(lj_context* ctx, const lj_params* p) {
    if (p->predicate == LJ_PRED_CUDALJ)
        return fail(ctx, LJ_ERR_INVALID_ARG, "the CUDALJ predicate has no cutoff; the LBVH needs one");
    return LJ_OK;
}

}  // namespace lj

using namespace lj;

extern "C" {

int lj_lbvh_compute(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
                    float* fx, float* fy, float* fz, float* e, lj_result* result, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    LJ_TRY(make_geom(ctx, p, n, &g, &c));
    LJ_TRY(check_lbvh_params(ctx, p));
    if (n > 0 && (!x || !y || !z || !fx || !fy || !fz))
        return fail(ctx, LJ_ERR_INVALID_ARG, "positions and forces must not be NULL");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    {
        PhaseTimer tb(ctx, LJ_PHASE_BIN, s);
        LJ_TRY(lbvh_build(ctx, x, y, z, n, p, s));
    }
    lj_result* res = result;
    if (!res) {
        LJ_HIP(ctx, ctx->bin.counters.grow(sizeof(lj_result)));
        res = ctx->bin.counters.as<lj_result>();
    }
    fill_u64(&res->n_candidates, 2, 0ull, s);
    const unsigned nb = div_up(n > 0 ? n : 1, 256);
    LJ_HIP(ctx, ctx->bin.partials.grow(((size_t)nb + nb / 1024 + 1) * 2 * sizeof(double)));
    double* partials = ctx->bin.partials.as<double>();
    int* ovf = overflow_flag(ctx, n);
    const float h = query_half_width(p);
    const float4* leaf = ctx->bin.pos.as<float4>();
    const int2* child = ctx->lbvh.node_child.as<int2>();
    const float4* box = ctx->lbvh.node_box.as<float4>();
    const bool vir = !(p->flags & LJ_FLAG_NO_VIRIAL);
#define LJ_LB(PRED, PREC, VIR)                                                                                        \
    hipLaunchKernelGGL((k_lbvh_force<PRED, PREC, VIR>), dim3(nb), dim3(256), 0, s, leaf, child, box, (int)n, h, c, fx, \
                       fy, fz, e, partials, res, ovf)
#define LJ_LB_PREC(PRED)                                                                                    \
    switch (p->precision) {                                                                                 \
    case LJ_PREC_MIXED: if (vir) LJ_LB(PRED, LJ_PREC_MIXED, true); else LJ_LB(PRED, LJ_PREC_MIXED, false); break; \
    case LJ_PREC_FP32: if (vir) LJ_LB(PRED, LJ_PREC_FP32, true); else LJ_LB(PRED, LJ_PREC_FP32, false); break;    \
    default: if (vir) LJ_LB(PRED, LJ_PREC_FP32_COMP, true); else LJ_LB(PRED, LJ_PREC_FP32_COMP, false); break;    \
    }
    PhaseTimer tf(ctx, LJ_PHASE_FORCE, s);
    if (n > 0) {
        switch (p->predicate) {
        case LJ_PRED_AABB: LJ_LB_PREC(LJ_PRED_AABB); break;
        case LJ_PRED_CUSTOM: LJ_LB_PREC(LJ_PRED_CUSTOM); break;
        case LJ_PRED_SPHERE: LJ_LB_PREC(LJ_PRED_SPHERE); break;
        case LJ_PRED_TRIANGLE: LJ_LB_PREC(LJ_PRED_TRIANGLE); break;
        }
    } else {
        hipLaunchKernelGGL((k_lbvh_force<LJ_PRED_AABB, LJ_PREC_MIXED, true>), dim3(1), dim3(256), 0, s, leaf, child, box,
                           0, h, c, fx, fy, fz, e, partials, res, ovf);
    }
#undef LJ_LB_PREC
#undef LJ_LB
    finalize_energy(partials, nb, res, s);
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

int lj_lbvh_pairs(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
                  unsigned long long* row_offsets, uint32_t* cols, unsigned long long capacity,
                  unsigned long long* needed, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    LJ_TRY(make_geom(ctx, p, n, &g, &c));
    LJ_TRY(check_lbvh_params(ctx, p));
    if (!row_offsets || (n > 0 && (!x || !y || !z)) || (capacity > 0 && !cols))
        return fail(ctx, LJ_ERR_INVALID_ARG, "NULL pointer argument");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    LJ_TRY(lbvh_build(ctx, x, y, z, n, p, s));
    const unsigned nb = div_up(n > 0 ? n : 1, 256);
    LJ_HIP(ctx, ctx->bin.rows.grow(((size_t)n + 1) * sizeof(uint32_t)));
    uint32_t* rowcnt = ctx->bin.rows.as<uint32_t>();
    int* ovf = overflow_flag(ctx, n);
    const float h = query_half_width(p);
    const float4* leaf = ctx->bin.pos.as<float4>();
    const int2* child = ctx->lbvh.node_child.as<int2>();
    const float4* box = ctx->lbvh.node_box.as<float4>();
    if (n > 0) {
        switch (p->predicate) {
#define LJ_C(P) case P: hipLaunchKernelGGL(k_lbvh_count<P>, dim3(nb), dim3(256), 0, s, leaf, child, box, (int)n, h, c, rowcnt, ovf); break;
            LJ_C(LJ_PRED_AABB) LJ_C(LJ_PRED_CUSTOM) LJ_C(LJ_PRED_SPHERE) LJ_C(LJ_PRED_TRIANGLE)
#undef LJ_C
        }
    }
    LJ_TRY(scan_u32_to_u64(ctx, rowcnt, row_offsets, n, s));
    unsigned long long total = 0;
    int overflow = 0;
    LJ_HIP(ctx, hipMemcpyAsync(&total, row_offsets + n, sizeof total, hipMemcpyDeviceToHost, s));
    LJ_HIP(ctx, hipMemcpyAsync(&overflow, ovf, sizeof overflow, hipMemcpyDeviceToHost, s));
    LJ_HIP(ctx, hipStreamSynchronize(s));
    if (overflow) return fail(ctx, LJ_ERR_OVERFLOW, "LBVH deeper than the traversal stack (flag %d)", overflow);
    if (needed) *needed = total;
    if (total > capacity) return fail(ctx, LJ_ERR_CAPACITY, "pair list needs %llu entries, capacity is %llu", total, capacity);
    if (total == 0) return LJ_OK;
    switch (p->predicate) {
#define LJ_F(P) case P: hipLaunchKernelGGL(k_lbvh_fill<P>, dim3(nb), dim3(256), 0, s, leaf, child, box, (int)n, h, c, (const unsigned long long*)row_offsets, cols, ovf); break;
        LJ_F(LJ_PRED_AABB) LJ_F(LJ_PRED_CUSTOM) LJ_F(LJ_PRED_SPHERE) LJ_F(LJ_PRED_TRIANGLE)
#undef LJ_F
    }
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

}  // extern "C"
