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
//  1. 27-bit Morton codes of the positions in the params box, stable radix sort
//     (ties keep index order, so the sorted (code, index) keys are unique);
//  2. the hierarchy of Karras, "Maximizing Parallelism in the Construction of BVHs,
//     Octrees, and k-d Trees", HPG 2012 (one thread per internal node);
//  3. node bounds without synchronisation: an internal node covers a contiguous range
//     of leaves (Morton order), so its bounds are the min/max over that range, read
//     from twelve levels of aligned chunk bounds (4, 16, ..., 2^24 leaves) plus the
//     leaves at the range's ends.  (Merging child boxes bottom-up would need an
//     acquire/release handshake per node across workgroups; at device scope on a
//     multi-XCD part each one writes back and invalidates the L2.)
//  4. one thread per particle (in Morton order) walks the tree without a stack: every
//     node stores, in the spare lanes of its bounds, its left child and its rope (the
//     node that follows its subtree in depth-first order), so a step descends on
//     overlap and follows the rope otherwise.  Subtrees of at most kBucket leaves are
//     buckets: one box test, then a scan of their leaves, which are contiguous.  The
//     query box [p - rc', p + rc'] is tested against node point-bounds, then the
//     canonical predicate and the pair arithmetic of the cell grid run on each leaf.
// A node stores the bounds of its particle positions; testing the query box of
// half-width rc' = rc (1 + 2^-18) against them is the point-in-AABB test of the
// reference with the rounding of p +- rc moved to the safe side: every pair the
// predicate can accept has |d_axis| <= rc (1 + 2^-20).
#include <cmath>

#include "lj_internal.h"
#include "lj_pair.h"
#include "lj_sort.h"

namespace lj {

namespace {

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

// Morton precision: 2^kMortonBits steps per axis of the box.  The codes only order the
// leaves (node bounds come from the positions, equal codes are split by index), and
// 27-bit codes sort in three 9-bit radix passes where 30-bit ones take four.
constexpr int kMortonBits = 9;

__device__ inline uint32_t quant(float x, float lo, float inv) {
    float q = (x - lo) * inv;
    q = fminf(fmaxf(q, 0.0f), (float)((1 << kMortonBits) - 1));
    return (uint32_t)q;
}

// 3 kMortonBits-bit Morton code of a position.
struct MortonKeyOp {
    float lo[3], inv[3];
    __device__ uint32_t operator()(float x, float y, float z, uint32_t) const {
        return (spread10(quant(x, lo[0], inv[0])) << 2) | (spread10(quant(y, lo[1], inv[1])) << 1) |
               spread10(quant(z, lo[2], inv[2]));
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

// Node numbering: internal nodes 0..n-2 (0 is the root), leaf k is node n-1+k (n <=
// 2^30, so every id fits 31 bits).  A child or rope reference with bit 31 set is a
// bucket: an internal node covering at most kBucket leaves, scanned instead of walked.
// Measured on 4,194,304 clustered particles, rc = 0.005, traversal + forces per step,
// buckets of B leaves scanned G at a time: B/G 8/2 2.82 ms, 12/2 2.66, 16/2 2.61, 24/2
// 2.65, 16/1 2.75, 12/4 2.61, 16/4 2.56, 24/4 2.61, 32/4 2.68.  Smaller buckets scan
// fewer leaves outside the query box (16: 173 leaves per query for 53 inside it) and
// walk more nodes (16: 111 steps per query, 24: 97).
constexpr int kBucket = 16;
constexpr int kCollapsed = (int)0x80000000u;
constexpr int kSentinel = 0x7fffffff;  // the rope of the last subtree: end of the walk

// Karras' construction.  Every split position gamma in [0, n-2] belongs to exactly
// one internal node, whose right child starts at leaf gamma + 1; that right child is
// therefore the rope of every subtree that ends at leaf gamma: rope_by_end[gamma].
// A walk reads the box of the root, of the nodes above kBucket leaves and of bucket
// roots, never of a node inside a bucket (about 7 in 8 nodes).  Every internal node but
// the root is the child of exactly one node, which marks it when it is a bucket root
// (so the marks need no clearing); k_node_bounds computes the boxes of the marked, the
// large and the root nodes only.
__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ code, int n, int2* __restrict__ child,
                                                int2* __restrict__ range, int* __restrict__ rope_by_end,
                                                uint8_t* __restrict__ bucket_root) {
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
    int left = (lo == gamma) ? (n - 1 + gamma) : gamma;
    int right = (hi == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    const bool big = hi - lo + 1 > kBucket;
    if (left < n - 1) {
        bucket_root[left] = big && gamma - lo + 1 <= kBucket;
        if (gamma - lo + 1 <= kBucket) left |= kCollapsed;
    }
    if (right < n - 1) {
        bucket_root[right] = big && hi - gamma <= kBucket;
        if (hi - gamma <= kBucket) right |= kCollapsed;
    }
    child[i] = make_int2(left, right);
    range[i] = make_int2(lo, hi);
    rope_by_end[gamma] = right;
}

__global__ __launch_bounds__(256) void k_fill_i32(int* __restrict__ p, uint32_t n, int v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

constexpr int kChunkBits = 2;  // a chunk of level l + 1 holds 4 chunks of level l
constexpr int kChunk = 1 << kChunkBits;
constexpr int kLevels = 24 / kChunkBits;  // chunks of 4, 16, ..., 2^24 leaves
static_assert(kLevels * kChunkBits == 24, "the largest chunk holds 2^24 leaves");
static_assert(kChunk == 4 && kLevels % 4 == 0, "k_chunk_bounds builds four levels of 4-way chunks per launch");

// First chunk pair of each level in the chunk array (level l holds ceil(n / 4^(l+1))).
struct ChunkLevels {
    uint32_t at[kLevels];
};

struct Bounds {
    float lo[3], hi[3];
    __device__ void empty() {
        lo[0] = lo[1] = lo[2] = INFINITY;
        hi[0] = hi[1] = hi[2] = -INFINITY;
    }
    __device__ void add(float4 l, float4 h) {
        lo[0] = fminf(lo[0], l.x);
        lo[1] = fminf(lo[1], l.y);
        lo[2] = fminf(lo[2], l.z);
        hi[0] = fmaxf(hi[0], h.x);
        hi[1] = fmaxf(hi[1], h.y);
        hi[2] = fmaxf(hi[2], h.z);
    }
};

// Bounds of aligned chunks, four levels per launch: a block of 256 threads takes 256
// consecutive inputs (leaf positions when `leaves`, else (lo, hi) pairs of the level
// below) and writes the chunks of 4, 16, 64 and 256 of them that it holds to out[0..3]
// (level L + j has cnt[j] chunks; a partial last chunk takes what exists).  Groups of
// 4, 16 and 64 lanes reduce with shuffles, the block's four waves through LDS.
struct ChunkOut {
    float4* out[4];
    uint32_t cnt[4];
};

__global__ __launch_bounds__(256) void k_chunk_bounds(const float4* __restrict__ in, bool leaves, uint32_t count,
                                                      ChunkOut o) {
    __shared__ float4 wlo[4], whi[4];
    const uint32_t e = blockIdx.x * 256u + threadIdx.x;  // input element
    Bounds b;
    b.empty();
    if (e < count) {
        if (leaves) b.add(in[e], in[e]);
        else b.add(in[2 * e], in[2 * e + 1]);
    }
    auto reduce = [&](int from, int to) {  // xor distances from .. to (powers of two)
        for (int d = from; d >= to; d >>= 1)
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = fminf(b.lo[a], __shfl_xor(b.lo[a], d, kWaveSize));
                b.hi[a] = fmaxf(b.hi[a], __shfl_xor(b.hi[a], d, kWaveSize));
            }
    };
    auto put = [&](int j, uint32_t c) {
        if (c < o.cnt[j]) {
            o.out[j][2 * c] = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.f);
            o.out[j][2 * c + 1] = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.f);
        }
    };
    reduce(2, 1);
    if (e % 4 == 0) put(0, e / 4);
    reduce(8, 4);
    if (e % 16 == 0) put(1, e / 16);
    reduce(32, 16);
    if (e % 64 == 0) {
        put(2, e / 64);
        wlo[threadIdx.x / 64] = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.f);
        whi[threadIdx.x / 64] = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.f);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        b.empty();
        for (int w = 0; w < 4; ++w) b.add(wlo[w], whi[w]);
        put(3, blockIdx.x);
    }
}

// Bounds of internal node i: min/max of the leaf positions in its range, from the
// largest aligned chunks inside the range and single leaves at its ends (at most
// kChunk - 1 elements per level on each side; narrow chunks keep that short).  Level l
// (0 .. kLevels-1, chunks of kChunk^(l + 1) leaves) starts at pair lv.at[l] of chunks.
// The .w lanes carry the walk: lo.w = the left child, hi.w = the rope (int bits).
__device__ inline void node_box(const float4* __restrict__ leaf, const int2* __restrict__ range,
                                const int2* __restrict__ child, const int* __restrict__ rope_by_end,
                                const float4* __restrict__ chunks, const ChunkLevels& lv, int i,
                                float4* __restrict__ box) {
    const int2 r = range[i];
    uint32_t a = (uint32_t)r.x;
    const uint32_t b = (uint32_t)r.y + 1u;  // exclusive, <= n
    Bounds bb;
    bb.empty();
    // element e of level l: a leaf (l = -1) or chunk e of level l
    auto add = [&](int l, uint32_t e) {
        if (l < 0) {
            const float4 p = leaf[e];
            bb.add(p, p);
        } else {
            const float4* c = chunks + 2 * ((size_t)lv.at[l] + e);
            bb.add(c[0], c[1]);
        }
    };
    // Climb: while a whole chunk of the next level lies inside [a, b), consume the
    // elements of the current level up to that chunk's start (a stays aligned to the
    // current level's size).
    int l = -1;
    while (l + 1 < kLevels) {
        const uint32_t size = 1u << (kChunkBits * (l + 1));
        const uint32_t up = 1u << (kChunkBits * (l + 2));  // l + 2 <= kLevels: at most 2^24
        const uint32_t start = (a + up - 1u) / up * up;     // a, b < 2^31: no overflow
        if (start + up > b) break;
        for (; a < start; a += size) add(l, a / size);
        ++l;
    }
    // Descend: whole elements of each level from the largest down to single leaves.
    for (; l >= -1; --l) {
        const uint32_t size = 1u << (kChunkBits * (l + 1));
        for (; a + size <= b; a += size) add(l, a / size);
    }
    box[2 * i] = make_float4(bb.lo[0], bb.lo[1], bb.lo[2], __int_as_float(child[i].x));
    box[2 * i + 1] = make_float4(bb.hi[0], bb.hi[1], bb.hi[2], __int_as_float(rope_by_end[r.y]));
}

// The boxes of the nodes a walk reads.  Each wave takes kNodesPerWave consecutive
// nodes, lists in LDS those whose box a walk reads, and computes them 64 at a time, so
// its lanes stay busy although only about one node in eight needs a box (in node order
// about as many lanes would idle).
constexpr int kNodeBoundsBlock = 256;
constexpr int kNodesPerWave = 512;
__global__ __launch_bounds__(kNodeBoundsBlock) void k_node_bounds(const float4* __restrict__ leaf, int n,
                                                     const int2* __restrict__ range, const int2* __restrict__ child,
                                                     const int* __restrict__ rope_by_end,
                                                     const float4* __restrict__ chunks, ChunkLevels lv,
                                                     const uint8_t* __restrict__ bucket_root,
                                                     float4* __restrict__ box) {
    __shared__ int todo[kNodeBoundsBlock / kWaveSize][kNodesPerWave];
    int* const list = todo[threadIdx.x / kWaveSize];
    const int first = (int)((blockIdx.x * (kNodeBoundsBlock / kWaveSize) + threadIdx.x / kWaveSize) * kNodesPerWave);
    // Every node's range and mark are loaded before any is tested (a node past n - 2
    // reads node 0's, n >= 2, and is not listed): tested as they came in, the loads went
    // out one group of 64 nodes at a time.  The root's mark is never written (it has no
    // parent); it is read here but not used.
    constexpr int kGroups = kNodesPerWave / kWaveSize;
    int2 rr[kGroups];
    uint8_t mark[kGroups];
#pragma unroll
    for (int u = 0; u < kGroups; ++u) {
        const int i = first + u * kWaveSize + (int)__lane_id();
        const int ic = i < n - 1 ? i : 0;
        rr[u] = range[ic];
        mark[u] = bucket_root[ic];
    }
    unsigned cnt = 0;  // wave-uniform
#pragma unroll
    for (int u = 0; u < kGroups; ++u) {
        const int i = first + u * kWaveSize + (int)__lane_id();
        const bool need = i < n - 1 && (i == 0 || rr[u].y - rr[u].x + 1 > kBucket || mark[u]);
        const uint64_t m = __ballot(need);
        if (need) list[cnt + sortk::rank_below(m)] = i;
        cnt += (unsigned)__popcll(m);
    }
    for (unsigned j = __lane_id(); j < cnt; j += kWaveSize)
        node_box(leaf, range, child, rope_by_end, chunks, lv, list[j], box);
}

// Does the query box [p - h, p + h] meet node bounds [lo, hi]?  Per axis, the gap
// max(lo - p, p - hi) is how far p lies outside [lo, hi] (negative inside); the boxes
// meet when no gap exceeds h.  Differences in FP32 are monotone in their operands, so
// the test never rejects a point within h.  For finite values this is exactly the six
// one-sided tests.  fmaxf drops a NaN difference when the other operand is a number, so
// the test then accepts more nodes, never fewer, and the leaf test decides; only when
// all six differences are NaN (a NaN query, or bounds no finite leaf made) does it
// reject, and then no leaf passes the leaf test either.  The max form is 6 subtractions,
// 2 three-way maxima and one compare, where six compares joined by && become a nest of
// divergent branches.
__device__ inline bool overlaps(float4 p, float h, float4 lo, float4 hi) {
    const float gx = fmaxf(lo.x - p.x, p.x - hi.x);
    const float gy = fmaxf(lo.y - p.y, p.y - hi.y);
    const float gz = fmaxf(lo.z - p.z, p.z - hi.z);
    return fmaxf(fmaxf(gx, gy), gz) <= h;
}

// Visits every leaf whose position lies within the query box of leaf k, calling
// f(slot, leaf position) for each one except k itself.  Depth-first, without a stack:
// on overlap go to the left child, otherwise (and after a leaf or bucket) follow the
// rope.  The walk runs in two phases so that a wave does the same kind of work at the
// same time: (1) walk the internal nodes and list, in LDS, the buckets and lone leaves
// that meet the query box, until the walk ends or the list holds kList entries; (2)
// scan the listed leaves in list order; repeat while the walk is unfinished (it stops
// at a node and resumes there, having no stack).  Interleaving the two made every wave
// step pay for a bucket scan as soon as one lane needed one.  The leaves are visited
// in walk order either way.  Returns false if the walk exceeds `limit` steps (2n: a
// well-formed tree never needs more, but it is checked).  list: this lane's entries, at list[q * stride].
// The list is LDS (kList * kLbvhBlock ints per block), which bounds the waves per CU:
// 16 entries and 64-thread blocks measured fastest (list 8..64, blocks 64..256).
//
// The loops are written for the SIMT machine: a wave runs every path any of its lanes
// takes, so internal nodes and buckets share one step (load the two box halves, test,
// pick the next node, list a bucket that meets the box) with no branch on the kind of
// node, and a leaf test is three compares of |p_j - p| (fl(a - b) = -fl(b - a), so
// |d| <= h is exactly the two one-sided tests of the reference's point-in-box) joined
// by `&`.  Lone leaves (leaves outside every bucket: on C5, one step in a thousand)
// keep a branch of their own.
constexpr int kList = 16;
constexpr int kLbvhBlock = 64;
constexpr int kLeafGroup = 4;  // bucket leaves loaded and tested per step of the scan
static_assert(kLeafGroup - 1 <= kPosSpare, "unclamped leaf loads need spare entries");

template <class F>
__device__ inline bool traverse(const float4* __restrict__ leaf, const int2* __restrict__ range,
                                const float4* __restrict__ box, const int* __restrict__ rope_by_end, int n, int k,
                                float4 p, float h, uint32_t limit, unsigned long long& tests, int* __restrict__ list,
                                uint32_t stride, F&& f) {
    if (n == 1) return true;
    uint32_t in_box = 0;
    auto leaf_test = [&](int slot, const float4& pj) {
        const bool in = (fabsf(pj.x - p.x) <= h) & (fabsf(pj.y - p.y) <= h) & (fabsf(pj.z - p.z) <= h);
        in_box += in ? 1u : 0u;
        if (in & (slot != k)) f(slot, pj);
    };
    int node = n <= kBucket ? (0 | kCollapsed) : 0;  // a small tree is one bucket
    uint32_t steps = 0;
    bool ok = true;
    while (node != kSentinel) {
        // phase 1: entries are bucket nodes (< n - 1) or lone leaves (node ids >= n - 1)
        unsigned cnt = 0;
        while (node != kSentinel && cnt < (unsigned)kList) {
            if (++steps > limit) {
                ok = false;
                node = kSentinel;
                break;
            }
            const int b = node & 0x7fffffff;
            if (b >= n - 1) {  // a lone leaf
                list[(cnt++) * stride] = b;
                node = rope_by_end[b - (n - 1)];
                continue;
            }
            const float4 lo = box[2 * b], hi = box[2 * b + 1];
            const bool ov = overlaps(p, h, lo, hi);
            const bool bucket = node < 0;
            if (ov & bucket) list[(cnt++) * stride] = b;
            node = (ov & !bucket) ? __float_as_int(lo.w) : __float_as_int(hi.w);
        }
        // phase 2: a bucket's leaves in groups of kLeafGroup, loaded together (the leaf
        // array has spare entries past the last leaf, so the loads need no clamp)
        for (unsigned q = 0; q < cnt; ++q) {
            const int e = list[q * stride];
            int2 r;
            if (e >= n - 1) r = make_int2(e - (n - 1), e - (n - 1));
            else r = range[e];
            for (int slot = r.x; slot <= r.y; slot += kLeafGroup) {
                float4 pj[kLeafGroup];
#pragma unroll
                for (int u = 0; u < kLeafGroup; ++u) pj[u] = leaf[slot + u];
                leaf_test(slot, pj[0]);
#pragma unroll
                for (int u = 1; u < kLeafGroup; ++u)
                    if (slot + u <= r.y) leaf_test(slot + u, pj[u]);
            }
        }
    }
    tests += in_box;
    return ok;
}

// STAGED: outputs as one staged record per particle (large outputs), else ordinary
// stores.
template <int PRED, int PREC, bool VIR, bool STAGED>
__global__ __launch_bounds__(kLbvhBlock) void k_lbvh_force(const float4* __restrict__ leaf, const int2* __restrict__ range,
                                                    const int* __restrict__ rope_by_end, const float4* __restrict__ box, int n, float h, uint32_t limit, PairConst c,
                                                    float* __restrict__ fx, float* __restrict__ fy,
                                                    float* __restrict__ fz, float* __restrict__ e,
                                                    double* __restrict__ partials, lj_result* __restrict__ res,
                                                    int* __restrict__ overflow, float4* __restrict__ stage) {
    extern __shared__ int lds_list[];
    int* list = lds_list + threadIdx.x;
    // consecutive queries walk nearly the same nodes: keep them on one XCD (own L2)
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int k = (int)(blk * blockDim.x + threadIdx.x);
    PairSums<PREC, false, VIR> acc;
    unsigned long long np = 0, tests = 0;
    double esum = 0.0, vsum = 0.0;
    if (k < n) {
        const float4 pi = leaf[k];
        acc.set_i(pi);
        acc.allow_exact_diff(c, pi, true);  // over the lanes with a query
        const bool ok = traverse(leaf, range, box, rope_by_end, n, k, pi, h, limit, tests, list, blockDim.x, [&](int, float4 pj) {
            const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const float d2 = dist2(dx, dy, dz);
            const int mult = pair_mult<PRED>(c, pi, pj, d2);
            if (!mult) return;
            np += (unsigned long long)mult;
            acc.add(c, pj, dx, dy, dz, d2, mult);
        });
        if (!ok) atomicOr(overflow, 1);
        const uint32_t i = __float_as_uint(pi.w);
        const float ox = (float)(-c.e24 * (double)acc.ax.value());
        const float oy = (float)(-c.e24 * (double)acc.ay.value());
        const float oz = (float)(-c.e24 * (double)acc.az.value());
        const double ei = c.e4 * (double)acc.ae.value();
        if constexpr (STAGED) {  // large outputs at shuffled indices: one staged record
            store_staged(&stage[i], ox, oy, oz, (float)ei);
        } else {
            fx[i] = ox;
            fy[i] = oy;
            fz[i] = oz;
            if (e) e[i] = (float)ei;
        }
        esum = ei;
        vsum = c.e24 * (double)acc.av.value();
    }
    block_sum2(esum, vsum);
    const unsigned long long bp = block_sum_u64(np);
    const unsigned long long bt = block_sum_u64(tests);
    if (threadIdx.x == 0) {
        partials[2 * blk] = esum;
        partials[2 * blk + 1] = vsum;
        if (bp) atomicAdd(&res->n_pairs, bp);
        if (bt) atomicAdd(&res->n_candidates, bt);
    }
}

template <int PRED>
__global__ __launch_bounds__(kLbvhBlock) void k_lbvh_count(const float4* __restrict__ leaf, const int2* __restrict__ range,
                                                    const int* __restrict__ rope_by_end, const float4* __restrict__ box, int n, float h, uint32_t limit, PairConst c,
                                                    uint32_t* __restrict__ rowcnt, int* __restrict__ overflow) {
    extern __shared__ int lds_list[];
    int* list = lds_list + threadIdx.x;
    const int k = (int)(xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    if (k >= n) return;
    const float4 pi = leaf[k];
    uint32_t cnt = 0;
    unsigned long long tests = 0;
    const bool ok = traverse(leaf, range, box, rope_by_end, n, k, pi, h, limit, tests, list, blockDim.x, [&](int, float4 pj) {
        if (pair_mult<PRED>(c, pi, pj, dist2(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z)) > 0) ++cnt;
    });
    if (!ok) atomicOr(overflow, 1);
    rowcnt[__float_as_uint(pi.w)] = cnt;
}

template <int PRED>
__global__ __launch_bounds__(kLbvhBlock) void k_lbvh_fill(const float4* __restrict__ leaf, const int2* __restrict__ range,
                                                   const int* __restrict__ rope_by_end, const float4* __restrict__ box, int n, float h, uint32_t limit, PairConst c,
                                                   const unsigned long long* __restrict__ rows,
                                                   uint32_t* __restrict__ cols, int* __restrict__ overflow) {
    extern __shared__ int lds_list[];
    int* list = lds_list + threadIdx.x;
    const int k = (int)(xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    if (k >= n) return;
    const float4 pi = leaf[k];
    const uint32_t i = __float_as_uint(pi.w);
    const unsigned long long r0 = rows[i], r1 = rows[i + 1];
    unsigned long long w = r0;
    unsigned long long tests = 0;
    const bool ok = traverse(leaf, range, box, rope_by_end, n, k, pi, h, limit, tests, list, blockDim.x, [&](int, float4 pj) {
        if (w < r1 && pair_mult<PRED>(c, pi, pj, dist2(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z)) > 0)
            cols[w++] = __float_as_uint(pj.w);
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

// Marks the results of a force call whose walk tripped its step guard: n_pairs =
// UINT64_MAX and NaN totals, so a caller that reads the results sees the failure
// without the call synchronising.
__global__ void k_poison_on_overflow(const int* __restrict__ overflow, lj_result* __restrict__ res) {
    if (threadIdx.x == 0 && *overflow) {
        res->n_pairs = ~0ull;
        res->e_total = __builtin_nan("");
        res->virial = __builtin_nan("");
    }
}

}  // namespace

// Step guard of every walk: 2n steps, or one step under LJ_FLAG_WALK_GUARD_TEST (the
// host-side way to exercise the guard's error path; every kernel still ends normally).
static uint32_t walk_limit(const lj_params* p, int64_t n) {
    return (p->flags & LJ_FLAG_WALK_GUARD_TEST) ? 1u : 2u * (uint32_t)(n > 0 ? n : 1);
}

// The traversal's overflow flag.
static int* overflow_flag(lj_context* ctx, int64_t) { return ctx->lbvh.flags.as<int>(); }

// Builds the tree over n particles into ctx->lbvh.  h is the query half-width.
static int lbvh_build(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
                      hipStream_t s) {
    LbvhWork& L = ctx->lbvh;
    BinWork& B = ctx->bin;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    LJ_HIP(ctx, L.node_child.grow(nn * sizeof(int2)));
    LJ_HIP(ctx, L.bucket_root.grow(nn));
    LJ_HIP(ctx, L.range.grow(nn * sizeof(int2)));
    LJ_HIP(ctx, L.rope.grow(nn * sizeof(int)));
    LJ_HIP(ctx, L.flags.grow(64));
    LJ_HIP(ctx, L.node_box.grow(nn * 2 * sizeof(float4)));
    // chunk bounds: level l holds ceil(n / kChunk^(l + 1)) (lo, hi) pairs, levels back to back
    ChunkLevels lv;
    uint64_t pairs = 0;
    {
        uint64_t cnt = (uint64_t)nn;
        for (int l = 0; l < kLevels; ++l) {
            cnt = div_up(cnt, kChunk);
            lv.at[l] = (uint32_t)pairs;
            pairs += cnt;
        }
    }
    LJ_HIP(ctx, L.chunks.grow((size_t)pairs * 2 * sizeof(float4)));
    MortonKeyOp op;
    for (int d = 0; d < 3; ++d) {
        op.lo[d] = p->box_lo[d];
        op.inv[d] = (float)(1 << kMortonBits) / (p->box_hi[d] - p->box_lo[d]);
    }
    // the sorted payload is the leaf array: positions in Morton order, w = original index
    LJ_TRY(sort_particles(ctx, x, y, z, n, op, 3 * kMortonBits, s));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, overflow_flag(ctx, n), 1u, 0);
    if (n == 0) return LJ_OK;
    const uint32_t* code = B.sorted_keys();
    int2* range = L.range.as<int2>();
    int* rope = L.rope.as<int>();
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, rope + (n - 1), 1u, kSentinel);
    if (n > 1)
        hipLaunchKernelGGL(k_karras, dim3(div_up(n - 1, 256)), dim3(256), 0, s, code, (int)n, L.node_child.as<int2>(),
                           range, rope, L.bucket_root.as<uint8_t>());
    // chunk bounds, four levels per launch (levels 0-3 from the leaves, 4-7 from level 3,
    // 8-11 from level 7)
    const float4* leaves = B.pos.as<float4>();
    {
        uint64_t cnt = (uint64_t)n;
        const float4* in = leaves;
        for (int l = 0; l < kLevels; l += 4) {
            ChunkOut o;
            uint64_t c = cnt;
            for (int j = 0; j < 4; ++j) {
                c = div_up(c, kChunk);
                o.out[j] = L.chunks.as<float4>() + 2 * (size_t)lv.at[l + j];
                o.cnt[j] = (uint32_t)c;
            }
            hipLaunchKernelGGL(k_chunk_bounds, dim3((unsigned)div_up(cnt, 256)), dim3(256), 0, s, in, l == 0,
                               (uint32_t)cnt, o);
            in = o.out[3];
            cnt = o.cnt[3];
        }
    }
    if (n > 1)
        hipLaunchKernelGGL(k_node_bounds, dim3(div_up(n - 1, kNodeBoundsBlock / kWaveSize * kNodesPerWave)),
                           dim3(kNodeBoundsBlock), 0, s, leaves, (int)n,
                           (const int2*)range, (const int2*)L.node_child.as<int2>(), (const int*)rope,
                           (const float4*)L.chunks.as<float4>(), lv, (const uint8_t*)L.bucket_root.as<uint8_t>(),
                           L.node_box.as<float4>());
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

// Query half-width.  Every pair a cutoff predicate accepts has |d_axis| <= rc (1 +
// 2^-20); the TRIANGLE ray model measures the cutoff to the centre rebuilt from the
// panel corners, which can lie a few ulp of the coordinates off the particle, so its
// boxes grow by 2^-20 of the largest coordinate magnitude of the box.
static float query_half_width(const lj_params* p) {
    float h = p->rc * (1.0f + 0x1p-18f);
    if (p->predicate == LJ_PRED_TRIANGLE) {
        float mag = 0.0f;
        for (int d = 0; d < 3; ++d) mag = std::fmax(mag, std::fmax(std::fabs(p->box_lo[d]), std::fabs(p->box_hi[d])));
        h += mag * 0x1p-20f;
    }
    return h;
}

static int check_lbvh_params(lj_context* ctx, const lj_params* p, int64_t n) {
    if (p->predicate == LJ_PRED_CUDALJ)
        return fail(ctx, LJ_ERR_INVALID_ARG, "the CUDALJ predicate has no cutoff; the LBVH needs one");
    if (n > (int64_t)1 << 30)
        return fail(ctx, LJ_ERR_INVALID_ARG, "the LBVH takes at most 2^30 particles (node ids are 31-bit)");
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
    LJ_TRY(check_lbvh_params(ctx, p, n));
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
    const unsigned nb = div_up(n > 0 ? n : 1, kLbvhBlock);
    const size_t lds = (size_t)kList * kLbvhBlock * sizeof(int);
    LJ_HIP(ctx, ctx->bin.partials.grow(((size_t)nb + nb / 1024 + 1) * 2 * sizeof(double)));
    double* partials = ctx->bin.partials.as<double>();
    int* ovf = overflow_flag(ctx, n);
    const float h = query_half_width(p);
    const uint32_t lim = walk_limit(p, n);
    const float4* leaf = ctx->bin.pos.as<float4>();
    const int2* range = ctx->lbvh.range.as<int2>();
    const int* rope = ctx->lbvh.rope.as<int>();
    const float4* box = ctx->lbvh.node_box.as<float4>();
    const bool vir = !(p->flags & LJ_FLAG_NO_VIRIAL);
    // large outputs at shuffled indices go through the staging array
    const bool nt = non_temporal_outputs(n);
    float4* stage = nullptr;
    if (nt) {
        LJ_HIP(ctx, ctx->bin.stage.grow((size_t)n * sizeof(float4)));
        stage = ctx->bin.stage.as<float4>();
    }
#define LJ_LBN(PRED, PREC, VIR, NT)                                                                                  \
    hipLaunchKernelGGL((k_lbvh_force<PRED, PREC, VIR, NT>), dim3(nb), dim3(kLbvhBlock), lds, s, leaf, range, rope, box, \
                       (int)n, h, lim, c, fx, fy, fz, e, partials, res, ovf, stage)
#define LJ_LB(PRED, PREC, VIR)                  \
    do {                                        \
        if (nt) LJ_LBN(PRED, PREC, VIR, true);  \
        else LJ_LBN(PRED, PREC, VIR, false);    \
    } while (0)
#define LJ_LB_PREC(PRED)                                                                                    \
    switch (p->precision) {                                                                                 \
    case LJ_PREC_MIXED: if (vir) LJ_LB(PRED, LJ_PREC_MIXED, true); else LJ_LB(PRED, LJ_PREC_MIXED, false); break; \
    case LJ_PREC_FP32: if (vir) LJ_LB(PRED, LJ_PREC_FP32, true); else LJ_LB(PRED, LJ_PREC_FP32, false); break;    \
    default: if (vir) LJ_LB(PRED, LJ_PREC_FP32_COMP, true); else LJ_LB(PRED, LJ_PREC_FP32_COMP, false); break;    \
    }
    {
        PhaseTimer tf(ctx, LJ_PHASE_FORCE, s);
        if (n > 0) {
            switch (p->predicate) {
            case LJ_PRED_AABB: LJ_LB_PREC(LJ_PRED_AABB); break;
            case LJ_PRED_CUSTOM: LJ_LB_PREC(LJ_PRED_CUSTOM); break;
            case LJ_PRED_SPHERE: LJ_LB_PREC(LJ_PRED_SPHERE); break;
            case LJ_PRED_TRIANGLE: LJ_LB_PREC(LJ_PRED_TRIANGLE); break;
            }
        } else {
            hipLaunchKernelGGL((k_lbvh_force<LJ_PRED_AABB, LJ_PREC_MIXED, true, false>), dim3(1), dim3(kLbvhBlock), lds,
                               s, leaf, range, rope, box, 0, h, lim, c, fx, fy, fz, e, partials, res, ovf, stage);
        }
    }
#undef LJ_LB_PREC
#undef LJ_LB
#undef LJ_LBN
    PhaseTimer to(ctx, LJ_PHASE_OTHER, s);
    if (stage) split_outputs(stage, (uint32_t)n, fx, fy, fz, e, s);
    finalize_energy(partials, nb, res, s);
    hipLaunchKernelGGL(k_poison_on_overflow, dim3(1), dim3(64), 0, s, (const int*)ovf, res);
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
    LJ_TRY(check_lbvh_params(ctx, p, n));
    if (!row_offsets || (n > 0 && (!x || !y || !z)) || (capacity > 0 && !cols))
        return fail(ctx, LJ_ERR_INVALID_ARG, "NULL pointer argument");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    LJ_TRY(lbvh_build(ctx, x, y, z, n, p, s));
    const unsigned nb = div_up(n > 0 ? n : 1, kLbvhBlock);
    const size_t lds = (size_t)kList * kLbvhBlock * sizeof(int);
    LJ_HIP(ctx, ctx->bin.rows.grow(((size_t)n + 1) * sizeof(uint32_t)));
    uint32_t* rowcnt = ctx->bin.rows.as<uint32_t>();
    int* ovf = overflow_flag(ctx, n);
    const float h = query_half_width(p);
    const uint32_t lim = walk_limit(p, n);
    const float4* leaf = ctx->bin.pos.as<float4>();
    const int2* range = ctx->lbvh.range.as<int2>();
    const int* rope = ctx->lbvh.rope.as<int>();
    const float4* box = ctx->lbvh.node_box.as<float4>();
    if (n > 0) {
        switch (p->predicate) {
#define LJ_C(P) case P: hipLaunchKernelGGL(k_lbvh_count<P>, dim3(nb), dim3(kLbvhBlock), lds, s, leaf, range, rope, box, (int)n, h, lim, c, rowcnt, ovf); break;
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
    if (overflow) return fail(ctx, LJ_ERR_OVERFLOW, "LBVH walk exceeded its step guard (flag %d)", overflow);
    if (needed) *needed = total;
    if (total > capacity) return fail(ctx, LJ_ERR_CAPACITY, "pair list needs %llu entries, capacity is %llu", total, capacity);
    if (total == 0) return LJ_OK;
    switch (p->predicate) {
#define LJ_F(P) case P: hipLaunchKernelGGL(k_lbvh_fill<P>, dim3(nb), dim3(kLbvhBlock), lds, s, leaf, range, rope, box, (int)n, h, lim, c, (const unsigned long long*)row_offsets, cols, ovf); break;
        LJ_F(LJ_PRED_AABB) LJ_F(LJ_PRED_CUSTOM) LJ_F(LJ_PRED_SPHERE) LJ_F(LJ_PRED_TRIANGLE)
#undef LJ_F
    }
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

}  // extern "C"
