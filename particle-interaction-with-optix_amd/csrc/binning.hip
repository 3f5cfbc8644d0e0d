// binning.hip -- cell binning on gfx950: stable LSD radix sort of (cell key, original
// index) carrying the positions, cell offsets from the sorted keys, device-wide scans,
// and the 27-cell candidate count.
//
// Reference: ComputeNbParticlePerCells (Cuda-version/cudaLJ.cu:345-383),
// PrefixFullV2 / BuildPartialPrefix_AllSize / ApplyPartialPrefixV3
// (Cuda-version/prefixsum.hpp:46-215), InitNewCells (cudaLJ.cu:428-439),
// MoveParticlesToNewCells (cudaLJ.cu:441-471), ComputeNbInteractions
// (cudaLJ.cu:255-297).
//
// Differences in design:
//  - the scan is reduce-then-scan with wave64 shuffles and no host round trip or
//    per-call allocation (the reference mallocs, recurses on the host and
//    synchronises at every level, prefixsum.hpp:195-214);
//  - the scatter is a stable radix sort, so the order inside a cell is the input
//    order and every output is deterministic (the reference hands out slots with
//    atomic tickets, cudaLJ.cu:462);
//  - the cell offsets come from the sorted keys in one pass that writes every cell
//    once, with no per-cell count array or scan (the reference counts with atomics
//    and prefix-sums the counts, cudaLJ.cu:345-383 and prefixsum.hpp);
//  - the candidate count is a closed form over cells, not a pass over particles.
#include "lj_internal.h"
#include "lj_sort.h"

namespace lj {

namespace {

constexpr int kWave = 64;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048 values per block

__device__ inline unsigned lane_id() { return __lane_id(); }

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- scan
template <class T>
__device__ inline T wave_incl_scan_t(T v) {
    const unsigned lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T o = __shfl_up(v, d, kWave);
        if (lane >= (unsigned)d) v += o;
    }
    return v;
}

// Block sums of kScanTile-value tiles (TI: input type, TS: sum type).
template <class TI, class TS>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const TI* __restrict__ in, uint32_t n,
                                                             TS* __restrict__ partials) {
    __shared__ TS wsum[kScanThreads / kWave];
    const uint32_t base = blockIdx.x * kScanTile;
    // every load issued before any is used: indices past n read the last value (n >= 1)
    // and count zero (a guarded load per item kept one in flight at a time)
    TI v[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t idx = base + k * kScanThreads + threadIdx.x;
        v[k] = in[idx < n ? idx : n - 1];
    }
    TS s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k * kScanThreads + threadIdx.x < n) s += (TS)v[k];
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) s += __shfl_xor(s, d, kWave);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        TS t = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) t += wsum[w];
        partials[blockIdx.x] = t;
    }
}

// Exclusive scan of one tile per block, offset by the scanned block partial.
// Thread t owns the kScanItems consecutive values [t*kScanItems, (t+1)*kScanItems).
// in may alias out when TI == TS.
template <class TI, class TS>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const TI* in, TS* out, uint32_t n,
                                                           const TS* __restrict__ block_off, TS* __restrict__ total) {
    __shared__ TS wsum[kScanThreads / kWave];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    TS v[kScanItems];
    TI raw[kScanItems];
    // A thread's eight consecutive 32-bit values as two 16-byte loads when they lie
    // inside n and in is 16-byte aligned (with eight 4-byte loads a wave's load touched
    // 16 lines); otherwise one clamped load each, as in k_scan_reduce.
    constexpr bool kVec = sizeof(TI) == 4 && kScanItems == 8;
    const bool vec = kVec && base + kScanItems <= n && ((uintptr_t)in & 15u) == 0;
    if (vec) {
        const uint4 a = *reinterpret_cast<const uint4*>(in + base);
        const uint4 b = *reinterpret_cast<const uint4*>(in + base + 4);
        const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) raw[k] = (TI)w[k];
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) raw[k] = in[base + k < n ? base + k : n - 1];
    }
    TS s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? (TS)raw[k] : (TS)0;
        s += v[k];
    }
    const TS incl = wave_incl_scan_t(s);
    const unsigned w = threadIdx.x / kWave;
    if (lane_id() == kWave - 1) wsum[w] = incl;
    __syncthreads();
    TS woff = block_off ? block_off[blockIdx.x] : (TS)0;
    for (unsigned q = 0; q < w; ++q) woff += wsum[q];
    TS run = woff + incl - s;
    if (kVec && sizeof(TS) == 4 && vec && ((uintptr_t)out & 15u) == 0) {
        uint32_t o[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            o[k] = (uint32_t)run;
            run += v[k];
        }
        *reinterpret_cast<uint4*>(out + base) = make_uint4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<uint4*>(out + base + 4) = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < n) out[base + k] = run;
            run += v[k];
        }
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) *total = run;
}

// ---------------------------------------------------------------- cell offsets
// off[c] = number of sorted keys below c, c in [0, ncells], straight from the sorted
// keys: the first slot k of each run of equal keys writes k into the cells above the
// previous run's key up to its own key, and the last slot writes n into the cells above
// the last key.  Every cell is written exactly once, so no fill, per-cell counts or
// scan are needed.  A lane writes a gap of up to kGapLane cells itself; its wave writes
// a gap of up to kGapInline cells, 64 cells per store; a longer gap is cut into pieces
// of at most kGapPiece cells, appended to a list that k_fill_gaps spreads over whole
// waves (sparse grids: one lane never loops over a long empty stretch).
constexpr uint32_t kGapLane = 8;
constexpr uint32_t kGapInline = 256;
constexpr uint32_t kGapPiece = 4096;
constexpr unsigned kGapBlocks = 1024;

struct GapPiece {
    uint32_t a, b, v;  // off[a .. b) = v
};

// off[a .. b) = v for each lane's range: short ranges by the lane itself, medium ones
// by the whole wave one range at a time, longer ones cut into pieces appended to the
// list (one atomic per wave: lanes take consecutive entries from a wave-wide prefix
// sum).  Every lane of the wave calls this.
__device__ inline void fill_cells(uint32_t a, uint32_t b, uint32_t v, uint32_t* __restrict__ off,
                                  GapPiece* __restrict__ gaps, uint32_t* __restrict__ ngaps) {
    const uint32_t len = b - a;
    const bool piecewise = len > kGapInline;
    if (len <= kGapLane)
        for (uint32_t c = a; c < b; ++c) off[c] = v;
    for (uint64_t m = __ballot(len > kGapLane && !piecewise); m; m &= m - 1) {
        const int l = __builtin_ctzll(m);
        const uint32_t la = __builtin_amdgcn_readlane(a, l), lb = __builtin_amdgcn_readlane(b, l);
        const uint32_t lv = __builtin_amdgcn_readlane(v, l);
        for (uint32_t c = la + lane_id(); c < lb; c += kWave) off[c] = lv;
    }
    if (!__any(piecewise)) return;
    const uint32_t pieces = piecewise ? (len + kGapPiece - 1) / kGapPiece : 0u;
    const uint32_t incl = wave_incl_scan_t(pieces);
    uint32_t at = 0;
    if (lane_id() == kWave - 1) at = atomicAdd(ngaps, incl);
    at = __shfl(at, kWave - 1, kWave) + incl - pieces;
    for (uint32_t i = 0; i < pieces; ++i) {
        const uint32_t pa = a + i * kGapPiece;
        gaps[at + i] = GapPiece{pa, b - pa > kGapPiece ? pa + kGapPiece : b, v};
    }
}

// The cells a wave writes, [A, E), are contiguous (the runs starting at its slots
// cover (key[k - 1], key[k]] one after the other, in lane order).  With STAGE (grids
// of 1.5 or more cells per particle, chosen by the host), a wave whose range holds
// kStageMin to kStage cells -- about one per lane or more, so that lanes would loop
// over short gaps with scattered stores -- fills a copy in LDS (short gaps lane by
// lane, longer ones by the whole wave) and stores it with consecutive lanes on
// consecutive cells; otherwise the cells are written in place as fill_cells does.
// (On grids of about one cell per particle the stores are already consecutive, and
// the staging code alone made the kernel slower.)  Each workgroup takes kOffChunks
// consecutive 256-key chunks and loads all their keys first: one key per thread left
// the kernel waiting on its loads (C3 on 32M cells 80 -> 50 us, on 16M cells 52 -> 39
// us, the C4 batch 252 -> 182 us).
constexpr uint32_t kStage = 1024;
constexpr uint32_t kStageMin = kWave;
constexpr int kOffChunks = 4;

template <bool STAGE>
__global__ __launch_bounds__(256) void k_offsets(const uint32_t* __restrict__ skeys, uint32_t n, uint32_t ncells,
                                                 uint32_t* __restrict__ off, GapPiece* __restrict__ gaps,
                                                 uint32_t* __restrict__ ngaps) {
    __shared__ uint32_t stage[STAGE ? 256 / kWave : 1][STAGE ? kStage : 1];
    // all the workgroup's keys are loaded before any cell is written
    uint32_t kk[kOffChunks], pk[kOffChunks];
#pragma unroll
    for (int u = 0; u < kOffChunks; ++u) {
        const uint32_t k = (blockIdx.x * kOffChunks + u) * blockDim.x + threadIdx.x;
        kk[u] = k < n ? skeys[k] : 0u;
        pk[u] = (k < n && k > 0) ? skeys[k - 1] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kOffChunks; ++u) {
        const uint32_t k = (blockIdx.x * kOffChunks + u) * blockDim.x + threadIdx.x;
        uint32_t a = 0, b = 0, v = 0;  // the cells (previous key, key] take slot k
        const uint32_t key = kk[u];
        if (k < n) {
            if (k == 0) {
                b = key + 1u;
            } else if (pk[u] != key) {
                a = pk[u] + 1u;
                b = key + 1u;
                v = k;
            }
        }
        uint32_t A = 0u, E = 0u;
        if constexpr (STAGE) {
            const uint64_t runs = __ballot(b > a);
            if (runs) {
                A = __builtin_amdgcn_readlane(a, __builtin_ctzll(runs));
                E = __builtin_amdgcn_readlane(b, 63 - __builtin_clzll(runs));
            }
        }
        if (STAGE && E - A >= kStageMin && E - A <= kStage) {
            uint32_t* st = stage[threadIdx.x / kWave];
            const uint32_t len = b - a;
            if (len <= kGapLane)
                for (uint32_t c = a; c < b; ++c) st[c - A] = v;
            for (uint64_t m = __ballot(len > kGapLane); m; m &= m - 1) {
                const int l = __builtin_ctzll(m);
                const uint32_t la = __builtin_amdgcn_readlane(a, l), lb = __builtin_amdgcn_readlane(b, l);
                const uint32_t lv = __builtin_amdgcn_readlane(v, l);
                for (uint32_t c = la + lane_id(); c < lb; c += kWave) st[c - A] = lv;
            }
            for (uint32_t c = lane_id(); c < E - A; c += kWave) off[A + c] = st[c];
        } else {
            fill_cells(a, b, v, off, gaps, ngaps);
        }
        // the cells above the last key take n (the wave holding slot n - 1)
        if (__any(k == n - 1)) {
            const bool last = k == n - 1;
            fill_cells(last ? key + 1u : 0u, last ? ncells + 1u : 0u, n, off, gaps, ngaps);
        }
    }
}

// One wave per listed piece.
__global__ __launch_bounds__(256) void k_fill_gaps(const GapPiece* __restrict__ gaps,
                                                   const uint32_t* __restrict__ ngaps, uint32_t* __restrict__ off) {
    const uint32_t m = *ngaps;
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwaves = gridDim.x * blockDim.x / kWave;
    for (uint32_t e = wave; e < m; e += nwaves) {
        const GapPiece p = gaps[e];
        for (uint32_t c = p.a + lane; c < p.b; c += kWave) off[c] = p.v;
    }
}

__global__ __launch_bounds__(256) void k_perm_from_pos(const float4* __restrict__ pos, uint32_t n,
                                                       uint32_t* __restrict__ perm) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) perm[k] = __float_as_uint(pos[k].w);
}

// ---------------------------------------------------------------- candidates
__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// sum over cells of n_c * (S_c - 1), S_c = the clipped 27-cell count including c;
// equals the per-particle count of ComputeNbInteractions (cudaLJ.cu:278-292).  One
// wave per (y,z) row of cells at a time, lanes along x: the row's neighbour rows are
// wave-uniform, so a cell's 27-cell count is nine x-range differences of the offsets
// (coalesced across the lanes) and no per-cell index arithmetic.  Each workgroup
// writes its sum to part[blockIdx.x] and k_sum_parts adds them up: on sparse grids
// (millions of mostly empty cells) the walk is a chain of dependent loads, so it needs
// many waves in flight, and atomics of that many workgroups to one counter would
// serialise.
constexpr unsigned kCandBlocks = 8192;

__global__ __launch_bounds__(256) void k_candidates(const uint32_t* __restrict__ off, GridGeom g,
                                                    unsigned long long* __restrict__ part) {
    constexpr unsigned kWaves = 256 / kWave;
    __shared__ unsigned long long wsum[kWaves];
    const int32_t gx = g.g[0], gy = g.g[1], gz = g.g[2];
    const uint32_t nrows = (uint32_t)gy * (uint32_t)gz;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int32_t lane = (int32_t)lane_id();
    unsigned long long v = 0;
    for (uint32_t r = blockIdx.x * kWaves + wave; r < nrows; r += gridDim.x * kWaves) {
        const int32_t cy = (int32_t)(r % (uint32_t)gy), cz = (int32_t)(r / (uint32_t)gy);
        const int32_t z0 = cz > 0 ? cz - 1 : 0, z1 = cz < gz - 1 ? cz + 1 : gz - 1;
        const int32_t y0 = cy > 0 ? cy - 1 : 0, y1 = cy < gy - 1 ? cy + 1 : gy - 1;
        const uint32_t* home = off + (size_t)r * (size_t)gx;
        for (int32_t cx = lane; cx < gx; cx += kWave) {
            const uint32_t nh = home[cx + 1] - home[cx];
            if (!nh) continue;
            // x-cells x0 .. xe - 1 of each neighbour row (the counts of at most n
            // particles fit 32 bits)
            const int32_t x0 = cx > 0 ? cx - 1 : 0, xe = cx < gx - 1 ? cx + 2 : gx;
            uint32_t sum = 0;
            for (int32_t zz = z0; zz <= z1; ++zz)
                for (int32_t yy = y0; yy <= y1; ++yy) {
                    const uint32_t* rw = off + ((size_t)zz * (size_t)gy + (size_t)yy) * (size_t)gx;
                    sum += rw[xe] - rw[x0];
                }
            v += (unsigned long long)nh * (sum - 1);
        }
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    if (lane == 0) wsum[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (unsigned w = 0; w < kWaves; ++w) t += wsum[w];
        part[blockIdx.x] = t;
    }
}

// *out = sum of part[0 .. n-1] (one workgroup of kSumThreads threads, each adding
// independent loads).
constexpr unsigned kSumThreads = 1024;
__global__ __launch_bounds__(kSumThreads) void k_sum_parts(const unsigned long long* __restrict__ part, uint32_t n,
                                                           unsigned long long* __restrict__ out) {
    constexpr unsigned kWaves = kSumThreads / kWave;
    __shared__ unsigned long long wsum[kWaves];
    unsigned long long v[4] = {0, 0, 0, 0};
    uint32_t i = threadIdx.x;
    for (; i + 3 * kSumThreads < n; i += 4 * kSumThreads)
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] += part[i + u * kSumThreads];
    for (; i < n; i += kSumThreads) v[0] += part[i];
    unsigned long long t = (v[0] + v[1]) + (v[2] + v[3]);
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) t += __shfl_xor(t, d, kWave);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (unsigned w = 0; w < kWaves; ++w) s += wsum[w];
        *out = s;
    }
}

int key_bits(uint32_t ncells) {
    int b = 0;
    while (b < 32 && ((uint64_t)1 << b) < (uint64_t)ncells) ++b;
    return b;
}

}  // namespace

// Workspace layout of scan_tmp: level k partial sums back to back.
static size_t scan_levels(int64_t n, int64_t* sizes, int* nlev) {
    size_t total = 0;
    int l = 0;
    int64_t m = n;
    while (m > kScanTile) {
        m = (m + kScanTile - 1) / kScanTile;
        sizes[l++] = m;
        total += (size_t)m;
    }
    *nlev = l;
    return total;
}

template <class TI, class TS>
static void scan_rec(const TI* in, TS* out, int64_t n, TS* tmp, const int64_t* sizes, int lev, int nlev,
                     TS* total_dev, hipStream_t s) {
    if (lev == nlev) {  // fits in one block
        hipLaunchKernelGGL((k_scan_down<TI, TS>), dim3(1), dim3(kScanThreads), 0, s, in, out, (uint32_t)n,
                           (const TS*)nullptr, total_dev);
        return;
    }
    const unsigned nb = div_up(n, kScanTile);
    TS* part = tmp;
    hipLaunchKernelGGL((k_scan_reduce<TI, TS>), dim3(nb), dim3(kScanThreads), 0, s, in, (uint32_t)n, part);
    scan_rec<TS, TS>(part, part, sizes[lev], tmp + sizes[lev], sizes, lev + 1, nlev, (TS*)nullptr, s);
    hipLaunchKernelGGL((k_scan_down<TI, TS>), dim3(nb), dim3(kScanThreads), 0, s, in, out, (uint32_t)n,
                       (const TS*)part, total_dev);
}

template <class TI, class TS>
static int scan_any(lj_context* ctx, const TI* in, TS* out, int64_t n, TS* total_dev, hipStream_t s) {
    if (n <= 0) {
        if (total_dev) {
            if constexpr (sizeof(TS) == 8) fill_u64((unsigned long long*)total_dev, 1, 0ull, s);
            else fill_u32((uint32_t*)total_dev, 1, 0u, s);
        }
        LJ_HIP(ctx, hipGetLastError());
        return LJ_OK;
    }
    int64_t sizes[8];
    int nlev = 0;
    const size_t need = scan_levels(n, sizes, &nlev);
    LJ_HIP(ctx, ctx->bin.scan_tmp.grow((need + 1) * sizeof(TS)));
    scan_rec<TI, TS>(in, out, n, ctx->bin.scan_tmp.as<TS>(), sizes, 0, nlev, total_dev, s);
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

int scan_u32(lj_context* ctx, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev, hipStream_t s) {
    return scan_any<uint32_t, uint32_t>(ctx, in, out, n, total_dev, s);
}

int scan_u32_to_u64(lj_context* ctx, const uint32_t* in, unsigned long long* out, int64_t n, hipStream_t s) {
    return scan_any<uint32_t, unsigned long long>(ctx, in, out, n, out + (n > 0 ? n : 0), s);
}

int cell_offsets_from_sorted(lj_context* ctx, int64_t n, uint32_t ncells, hipStream_t s) {
    BinWork& w = ctx->bin;
    LJ_HIP(ctx, w.offsets.grow(((size_t)ncells + 1) * 4));
    uint32_t* off = w.offsets.as<uint32_t>();
    if (n <= 0) {
        fill_u32(off, (uint64_t)ncells + 1, 0u, s);
        LJ_HIP(ctx, hipGetLastError());
        return LJ_OK;
    }
    // gaps are disjoint and each piece list entry covers kGapPiece cells, or ends a gap
    // of more than kGapInline cells
    const size_t cells = (size_t)ncells + 1;
    const size_t cap = cells / kGapPiece + cells / kGapInline + 2;
    LJ_HIP(ctx, w.gaps.grow(sizeof(uint32_t) * 4 + cap * sizeof(GapPiece)));
    uint32_t* ngaps = w.gaps.as<uint32_t>();
    GapPiece* gaps = reinterpret_cast<GapPiece*>(w.gaps.as<uint8_t>() + sizeof(uint32_t) * 4);
    fill_u32(ngaps, 1, 0u, s);
    if (2 * (uint64_t)ncells >= 3 * (uint64_t)n)
        hipLaunchKernelGGL(k_offsets<true>, dim3(div_up(n, 256 * kOffChunks)), dim3(256), 0, s, w.sorted_keys(), (uint32_t)n, ncells, off,
                       gaps, ngaps);
    else
        hipLaunchKernelGGL(k_offsets<false>, dim3(div_up(n, 256 * kOffChunks)), dim3(256), 0, s, w.sorted_keys(), (uint32_t)n, ncells, off,
                       gaps, ngaps);
    const unsigned nb = (unsigned)(cap < kGapBlocks ? cap : kGapBlocks);
    hipLaunchKernelGGL(k_fill_gaps, dim3(nb), dim3(256), 0, s, (const GapPiece*)gaps, (const uint32_t*)ngaps, off);
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

int bin_particles(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
                  const GridGeom& g, hipStream_t s) {
    LJ_TRY(sort_particles(ctx, x, y, z, n, CellKeyOp{g}, key_bits(g.ncells), s));
    return cell_offsets_from_sorted(ctx, n, g.ncells, s);
}

int extract_perm(lj_context* ctx, int64_t n, uint32_t* perm, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_perm_from_pos, dim3(div_up(n, 256)), dim3(256), 0, s, ctx->bin.pos.as<float4>(),
                           (uint32_t)n, perm);
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

int count_candidates_from(lj_context* ctx, const GridGeom& g, const uint32_t* offsets, unsigned long long* out,
                          hipStream_t s) {
    const uint64_t nrows = (uint64_t)g.g[1] * (uint64_t)g.g[2];
    const uint64_t rb = (nrows + 3) / 4;  // four waves (rows) per workgroup
    const unsigned nb = (unsigned)(rb < kCandBlocks ? (rb > 0 ? rb : 1) : kCandBlocks);
    LJ_HIP(ctx, ctx->bin.cand.grow((size_t)nb * sizeof(unsigned long long)));
    unsigned long long* part = ctx->bin.cand.as<unsigned long long>();
    hipLaunchKernelGGL(k_candidates, dim3(nb), dim3(256), 0, s, offsets, g, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kSumThreads), 0, s, (const unsigned long long*)part, nb, out);
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

int count_candidates(lj_context* ctx, const GridGeom& g, unsigned long long* out, hipStream_t s) {
    return count_candidates_from(ctx, g, ctx->bin.offsets.as<uint32_t>(), out, s);
}

void fill_u64(unsigned long long* p, uint32_t n, unsigned long long v, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u64, dim3(div_up(n, 256)), dim3(256), 0, s, p, n, v);
}

void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(div_up((int64_t)n, 256)), dim3(256), 0, s, p, n, v);
}

}  // namespace lj
