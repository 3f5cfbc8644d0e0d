// lj_sort.h -- stable LSD radix sort of particles by a 32-bit key, carrying the payload
// (x, y, z, original index) through every pass.
//
// Digits are 8 bits; keys of 25 to 27 bits take one to three 9-bit digits in their
// last passes (three passes instead of four).  One pass over RB key bits (R = 2^RB
// digits) is:
//   histogram  per 2048-key tile, tile-major table hist[tile * R + d] (one contiguous
//              row per tile)
//   offsets    a digit-major exclusive scan of that table in three small kernels:
//              hist[tile * R + d] becomes the first output slot of digit d's keys in
//              that tile
//   scatter    each wave ranks its own contiguous run of keys stably with per-wave LDS
//              counters (no block barrier per round), the waves' counts are combined
//              once, the tile is staged in LDS in digit order, and every digit's run
//              goes to consecutive output slots, so stores are contiguous runs rather
//              than one scattered 20-byte record per lane.
// The first pass computes the keys from the positions with a key functor (cell key,
// batched cell key, Morton code) and reads x, y, z in input order, so no pass ever
// gathers positions through a permutation.
#pragma once

#include "lj_internal.h"

namespace lj {

namespace sortk {

constexpr int kWave = 64;
constexpr int kThreads = 256;  // histogram
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;  // 2048 keys per tile
// scatter: more, shorter waves per tile than the histogram, so that more waves share a
// CU while the tile's staging area (which bounds the workgroups per CU) stays the same
constexpr int kSThreads = 1024;
constexpr int kSItems = kTile / kSThreads;
constexpr int kSWaves = kSThreads / kWave;
constexpr int kWaveKeys = kSItems * kWave;  // the contiguous keys of one wave in a tile
static_assert(kSItems * kSThreads == kTile, "scatter shape");
constexpr uint32_t kMaxChunks = 256;       // row groups of the offset scan

typedef float vf4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned rank_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Key + payload of element idx, from positions (first pass: the key read from kin,
// where the first histogram stored it, or computed when kin is null) or from the
// previous pass.
template <bool FIRST, class KeyOp>
__device__ inline void load_elem(uint32_t idx, const float* __restrict__ x, const float* __restrict__ y,
                                 const float* __restrict__ z, const uint32_t* __restrict__ kin,
                                 const float4* __restrict__ pin, const KeyOp& op, uint32_t& key, float4& pay) {
    if constexpr (FIRST) {
        const float px = x[idx], py = y[idx], pz = z[idx];
        key = kin ? kin[idx] : op(px, py, pz, idx);
        pay = make_float4(px, py, pz, __uint_as_float(idx));
    } else {
        key = kin[idx];
        pay = pin[idx];
    }
}

// The first pass computes the keys and stores them in kstore (input order) for its
// scatter, which then need not compute them again.
template <int RB, bool FIRST, class KeyOp>
__global__ __launch_bounds__(kThreads) void k_hist(const float* __restrict__ x, const float* __restrict__ y,
                                                   const float* __restrict__ z, const uint32_t* __restrict__ kin,
                                                   uint32_t n, KeyOp op, int shift, uint32_t* __restrict__ hist,
                                                   uint32_t* __restrict__ kstore) {
    constexpr int kRadix = 1 << RB;
    __shared__ uint32_t h[kRadix];
    for (int d = threadIdx.x; d < kRadix; d += kThreads) h[d] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
    // Every load of the thread is issued before any key is used: items past n load the
    // last element instead (n >= 1) and are not counted, so no branch separates the
    // loads (with a guard or an LDS atomic between them a wave had one item's loads in
    // flight at a time).
    uint32_t key[kItems];
    if constexpr (FIRST) {
        float px[kItems], py[kItems], pz[kItems];
#pragma unroll
        for (int r = 0; r < kItems; ++r) {
            const uint32_t idx = base + r * kThreads + threadIdx.x;
            const uint32_t ic = idx < n ? idx : n - 1;
            px[r] = x[ic];
            py[r] = y[ic];
            pz[r] = z[ic];
        }
#pragma unroll
        for (int r = 0; r < kItems; ++r) {
            const uint32_t idx = base + r * kThreads + threadIdx.x;
            key[r] = op(px[r], py[r], pz[r], idx < n ? idx : n - 1);
            if (idx < n) kstore[idx] = key[r];
        }
    } else {
#pragma unroll
        for (int r = 0; r < kItems; ++r) {
            const uint32_t idx = base + r * kThreads + threadIdx.x;
            key[r] = kin[idx < n ? idx : n - 1];
        }
    }
#pragma unroll
    for (int r = 0; r < kItems; ++r)
        if (base + r * kThreads + threadIdx.x < n) atomicAdd(&h[(key[r] >> shift) & (kRadix - 1)], 1u);
    __syncthreads();
    for (int d = threadIdx.x; d < kRadix; d += kThreads) hist[(uint64_t)blockIdx.x * kRadix + d] = h[d];
}

// ---- digit-major exclusive scan of the tile-major table, rows grouped in chunks of
// `chunk` consecutive tiles (at most kMaxChunks chunks).
// part[c * R + d] = sum of digit d over the rows of chunk c (one thread per digit).
template <int RB>
__global__ __launch_bounds__(1 << RB) void k_dig_reduce(const uint32_t* __restrict__ hist, uint32_t ntiles,
                                                        uint32_t chunk, uint32_t* __restrict__ part) {
    constexpr int kRadix = 1 << RB;
    const uint32_t t0 = blockIdx.x * chunk;
    const uint32_t t1 = t0 + chunk < ntiles ? t0 + chunk : ntiles;
    uint32_t s = 0;
#pragma unroll 8
    for (uint32_t t = t0; t < t1; ++t) s += hist[(uint64_t)t * kRadix + threadIdx.x];
    part[blockIdx.x * kRadix + threadIdx.x] = s;
}

// The chunk sums part[c * R + d] become their exclusive prefix over the chunks, and
// dtot[d] digit d's total.  A workgroup takes 64 digits, one per lane; its 16 waves
// take 16 consecutive chunks each, loaded at once, and combine their sums in LDS.
// (One 1024-thread block owning every digit, each thread walking a share of the
// chunks, took 10 us for 8-bit digits and 17 us for 9-bit ones on 16.7M keys, and so
// did one wave per digit with its lanes on strided chunks.)
template <int RB>
__global__ __launch_bounds__(1024) void k_dig_scan(uint32_t* __restrict__ part, uint32_t nchunks,
                                                   uint32_t* __restrict__ dtot) {
    constexpr int kRadix = 1 << RB, kW = 1024 / kWave, kPer = kMaxChunks / kW;
    static_assert(kRadix % kWave == 0 && kPer * kW == kMaxChunks, "digit scan shape");
    __shared__ uint32_t wsum[kW][kWave];
    const unsigned lane = __lane_id(), w = threadIdx.x / kWave;
    const uint32_t d = blockIdx.x * kWave + lane;
    uint32_t v[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const uint32_t c = w * kPer + i;
        v[i] = c < nchunks ? part[c * kRadix + d] : 0u;
        sum += v[i];
    }
    wsum[w][lane] = sum;
    __syncthreads();
    uint32_t run = 0;
    for (unsigned k = 0; k < w; ++k) run += wsum[k][lane];
    if (w == kW - 1) dtot[d] = run + sum;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const uint32_t c = w * kPer + i;
        if (c < nchunks) part[c * kRadix + d] = run;
        run += v[i];
    }
}

// hist[t * R + d] <- first output slot of digit d's keys in tile t: the keys of smaller
// digits (an exclusive scan of the digit totals, redone by every block), plus digit
// d's keys in earlier chunks, plus those in the chunk's earlier tiles.
template <int RB>
__global__ __launch_bounds__(1 << RB) void k_dig_down(uint32_t* __restrict__ hist, uint32_t ntiles, uint32_t chunk,
                                                      const uint32_t* __restrict__ part,
                                                      const uint32_t* __restrict__ dtot) {
    constexpr int kRadix = 1 << RB;
    __shared__ uint32_t wsum[kRadix / kWave];
    const unsigned d = threadIdx.x, lane = __lane_id();
    const uint32_t tot = dtot[d];
    uint32_t incl = tot;
#pragma unroll
    for (int k = 1; k < kWave; k <<= 1) {
        const uint32_t o = __shfl_up(incl, k, kWave);
        if (lane >= (unsigned)k) incl += o;
    }
    if (lane == kWave - 1) wsum[d / kWave] = incl;
    __syncthreads();
    uint32_t run = incl - tot + part[blockIdx.x * kRadix + d];
    for (unsigned w = 0; w < d / kWave; ++w) run += wsum[w];
    const uint32_t t0 = blockIdx.x * chunk;
    const uint32_t t1 = t0 + chunk < ntiles ? t0 + chunk : ntiles;
#pragma unroll 8
    for (uint32_t t = t0; t < t1; ++t) {
        const uint64_t at = (uint64_t)t * kRadix + d;
        const uint32_t v = hist[at];
        hist[at] = run;
        run += v;
    }
}

// Tile t (logical, XCD-aware: neighbouring tiles write neighbouring output runs and
// share an L2) sorts its keys stably by digit and writes each digit's run.  Element j
// of the tile (input order) is handled by wave j / kWaveKeys, round
// (j % kWaveKeys) / 64, lane j % 64, so (wave, round, lane) is the input order.
template <int RB, bool FIRST, class KeyOp>
__global__ __launch_bounds__(kSThreads) void k_scatter(const float* __restrict__ x, const float* __restrict__ y,
                                                       const float* __restrict__ z, const uint32_t* __restrict__ kin,
                                                       const float4* __restrict__ pin, uint32_t n, KeyOp op, int shift,
                                                       const uint32_t* __restrict__ hist, uint32_t* __restrict__ kout,
                                                       float4* __restrict__ pout) {
    constexpr int kRadix = 1 << RB;
    static_assert(kRadix <= kSThreads, "one thread per digit combines the waves' counts");
    __shared__ uint32_t skey[kTile];
    __shared__ float4 spay[kTile];
    // per-wave digit counts, then per-wave first slots, as 16-bit halves of 32-bit words
    // (digits d and d ^ 1 share word d / 2; every value is at most kTile), so that three
    // workgroups of 8-bit digits fit in a CU's LDS instead of two
    __shared__ uint32_t wcnt[kSWaves][kRadix / 2];
    uint16_t* const wc16 = reinterpret_cast<uint16_t*>(&wcnt[0][0]);  // [wave * kRadix + d]
    static_assert(kTile < 65536, "16-bit counters");
    __shared__ uint32_t dstart[kRadix];         // first staging slot of each digit in this tile
    __shared__ uint32_t gbase[kRadix];          // first output slot of each digit for this tile
    __shared__ uint32_t wsum[kRadix / kWave];
    const unsigned tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t base = tile * kTile;
    const uint32_t count = n - base < (uint32_t)kTile ? n - base : (uint32_t)kTile;

    uint32_t key[kSItems];
    float4 pay[kSItems];
    uint32_t rank[kSItems];
    // Every element load is issued first, with no branch between them: a lane past the
    // tile's end loads the tile's last element (count >= 1) and is never staged, so its
    // values are defined.  (With guarded loads a wave had one element's loads in flight
    // at a time, and the digit table's row, loaded before them, was waited for first.)
    // The first pass reads the keys its histogram stored (kin is never null here).
#pragma unroll
    for (int r = 0; r < kSItems; ++r) {
        const uint32_t j = wave * kWaveKeys + r * kWave + lane;
        const uint32_t idx = base + (j < count ? j : count - 1);
        key[r] = kin[idx];
        if constexpr (FIRST) {
            pay[r] = make_float4(x[idx], y[idx], z[idx], __uint_as_float(idx));
        } else {
            // as a native vector: a float4 struct copied here went through a private
            // (scratch) copy, 48 bytes per lane with the keys and ranks
            const vf4 v = reinterpret_cast<const vf4*>(pin)[idx];
            pay[r] = make_float4(v.x, v.y, v.z, v.w);
        }
    }
#pragma unroll
    for (int q = 0; q < kRadix / 2 / kWave; ++q) wcnt[wave][q * kWave + lane] = 0;
    if (tid < (unsigned)kRadix) gbase[tid] = hist[(uint64_t)tile * kRadix + tid];
    // Stable ranks inside the wave: the lowest lane of each digit group adds the group's
    // size to the wave's counter with an LDS atomic (the wave's own counters, so no
    // other wave touches them) and hands the old value to its group.
#pragma unroll
    for (int r = 0; r < kSItems; ++r) {
        const bool valid = wave * kWaveKeys + r * kWave + lane < count;
        const uint32_t d = (key[r] >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const unsigned rk = rank_below(peers);
        uint32_t before = 0;
        if (valid && rk == 0) {
            const uint32_t half = 16u * (d & 1u);
            before = (atomicAdd(&wcnt[wave][d >> 1], (uint32_t)__popcll(peers) << half) >> half) & 0xffffu;
        }
        const int leader = valid ? __builtin_ctzll(peers) : (int)lane;
        before = __shfl(before, leader, kWave);
        rank[r] = before + rk;
    }
    __syncthreads();
    // threads 0..R-1, one per digit: staging start of the digit (exclusive scan over
    // digits of the tile's counts) and of each wave's part of it
    uint32_t c[kSWaves];
    uint32_t tot = 0, incl = 0;
    if (tid < (unsigned)kRadix) {
#pragma unroll
        for (int w = 0; w < kSWaves; ++w) {
            c[w] = wc16[w * kRadix + tid];
            tot += c[w];
        }
        incl = tot;
#pragma unroll
        for (int k = 1; k < kWave; k <<= 1) {
            const uint32_t o = __shfl_up(incl, k, kWave);
            if (lane >= (unsigned)k) incl += o;
        }
        if (lane == kWave - 1) wsum[wave] = incl;
    }
    __syncthreads();
    if (tid < (unsigned)kRadix) {
        uint32_t run = incl - tot;
        for (unsigned q = 0; q < wave; ++q) run += wsum[q];
        dstart[tid] = run;
#pragma unroll
        for (int w = 0; w < kSWaves; ++w) {
            wc16[w * kRadix + tid] = (uint16_t)run;
            run += c[w];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSItems; ++r) {
        if (wave * kWaveKeys + r * kWave + lane < count) {
            const uint32_t d = (key[r] >> shift) & (kRadix - 1);
            const uint32_t slot = wc16[wave * kRadix + d] + rank[r];
            skey[slot] = key[r];
            spay[slot] = pay[r];
        }
    }
    __syncthreads();
    // write the staged tile: consecutive slots of one digit go to consecutive outputs
#pragma unroll
    for (int r = 0; r < kSItems; ++r) {
        const uint32_t p = r * kSThreads + tid;
        if (p < count) {
            const uint32_t k = skey[p];
            const uint32_t d = (k >> shift) & (kRadix - 1);
            const uint32_t dst = gbase[d] + (p - dstart[d]);
            kout[dst] = k;
            pout[dst] = spay[p];
        }
    }
}

// No sorting needed (one cell or one particle): keys and payload in input order.
template <class KeyOp>
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ x, const float* __restrict__ y,
                                              const float* __restrict__ z, uint32_t n, KeyOp op,
                                              uint32_t* __restrict__ kout, float4* __restrict__ pout) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k;
    float4 p;
    load_elem<true>(i, x, y, z, nullptr, nullptr, op, k, p);
    kout[i] = k;
    pout[i] = p;
}

}  // namespace sortk

namespace sortk {

// Pass p of `passes` of a sort (n >= 2), on the RB key bits from `shift`.  The keys
// ping-pong between keys[0] and keys[1] (kcur holds the current one), the payload
// between pos and pos_alt so that the last pass writes ctx->bin.pos.  The digit table
// (hist) has room for the largest radix.
template <int RB, class KeyOp>
void sort_pass(BinWork& w, uint32_t* hist, const float* x, const float* y, const float* z, int64_t n,
               const KeyOp& op, int p, int passes, int shift, int& kcur, hipStream_t s) {
    constexpr int kRadix = 1 << RB;
    const unsigned ntiles = div_up(n, kTile);
    const uint32_t chunk = div_up(ntiles, kMaxChunks);
    const uint32_t nchunks = div_up(ntiles, chunk);
    uint32_t* part = hist + (size_t)ntiles * kRadix;  // the chunk partials after the table,
    uint32_t* dtot = part + (size_t)kMaxChunks * kRadix;  // then the digit totals
    float4* pbuf[2] = {w.pos.as<float4>(), w.pos_alt.as<float4>()};
    const int kdst = p == 0 ? 0 : kcur ^ 1;
    float4* pdst = pbuf[(passes - 1 - p) & 1];
    const float4* psrc = pbuf[(passes - p) & 1];
    if (p == 0) {
        // the keys in input order go to keys[0], which pass 0 does not otherwise use
        hipLaunchKernelGGL((k_hist<RB, true, KeyOp>), dim3(ntiles), dim3(kThreads), 0, s, x, y, z,
                           (const uint32_t*)nullptr, (uint32_t)n, op, shift, hist, w.keys[0].as<uint32_t>());
    } else {
        hipLaunchKernelGGL((k_hist<RB, false, KeyOp>), dim3(ntiles), dim3(kThreads), 0, s, x, y, z,
                           (const uint32_t*)w.keys[kcur].as<uint32_t>(), (uint32_t)n, op, shift, hist,
                           (uint32_t*)nullptr);
    }
    hipLaunchKernelGGL(k_dig_reduce<RB>, dim3(nchunks), dim3(kRadix), 0, s, (const uint32_t*)hist, ntiles, chunk,
                       part);
    hipLaunchKernelGGL(k_dig_scan<RB>, dim3(kRadix / kWave), dim3(1024), 0, s, part, nchunks, dtot);
    hipLaunchKernelGGL(k_dig_down<RB>, dim3(nchunks), dim3(kRadix), 0, s, hist, ntiles, chunk, (const uint32_t*)part,
                       (const uint32_t*)dtot);
    if (p == 0) {
        // pass 0 reads positions and the stored keys; it writes keys[1]
        hipLaunchKernelGGL((k_scatter<RB, true, KeyOp>), dim3(ntiles), dim3(kSThreads), 0, s, x, y, z,
                           (const uint32_t*)w.keys[0].as<uint32_t>(), (const float4*)nullptr, (uint32_t)n, op,
                           shift, (const uint32_t*)hist, w.keys[1].as<uint32_t>(), pdst);
        kcur = 1;
    } else {
        hipLaunchKernelGGL((k_scatter<RB, false, KeyOp>), dim3(ntiles), dim3(kSThreads), 0, s, x, y, z,
                           (const uint32_t*)w.keys[kcur].as<uint32_t>(), psrc, (uint32_t)n, op, shift,
                           (const uint32_t*)hist, w.keys[kdst].as<uint32_t>(), pdst);
        kcur = kdst;
    }
}

// The radix passes of a sort over `bits` key bits (n >= 2): 8-bit digits, with 9-bit
// ones in the last passes where that saves a pass (25 to 27 bits: 8+8+9, 8+9+9,
// 9+9+9; the 9-bit digit scan takes 17 us on 16.7M keys where the 8-bit one takes 10).
template <class KeyOp>
int sort_passes(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const KeyOp& op,
                int bits, hipStream_t s) {
    BinWork& w = ctx->bin;
    const unsigned ntiles = div_up(n, kTile);
    int nine = 0, passes = (bits + 7) / 8;
    if (bits > 24 && bits <= 27) {
        passes = 3;
        nine = bits - 24;  // passes of 9 bits, the last ones
    }
    LJ_HIP(ctx, w.hist.grow(((size_t)ntiles + kMaxChunks + 1) * (nine ? 512 : 256) * 4));
    uint32_t* hist = w.hist.as<uint32_t>();
    int kcur = 0, shift = 0;
    for (int p = 0; p < passes; ++p) {
        if (p >= passes - nine) {
            sort_pass<9>(w, hist, x, y, z, n, op, p, passes, shift, kcur, s);
            shift += 9;
        } else {
            sort_pass<8>(w, hist, x, y, z, n, op, p, passes, shift, kcur, s);
            shift += 8;
        }
    }
    w.cur = kcur;
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

}  // namespace sortk

// Sorts n particles by op's key (the low `bits` bits).  The sorted keys land in
// ctx->bin.keys[ctx->bin.cur] and the payload (x, y, z, bits of the original index)
// in ctx->bin.pos, which has room for n + kPosSpare entries (the spare ones are not
// written).
template <class KeyOp>
int sort_particles(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const KeyOp& op,
                   int bits, hipStream_t s) {
    using namespace sortk;
    BinWork& w = ctx->bin;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    LJ_HIP(ctx, w.keys[0].grow(nn * 4));
    LJ_HIP(ctx, w.keys[1].grow(nn * 4));
    // spare slots beyond the last particle: the force loops read groups of slots
    // (t .. t + kCandidates - 1) without clamping and use a slot only when it is in range
    LJ_HIP(ctx, w.pos.grow((nn + kPosSpare) * sizeof(float4)));
    LJ_HIP(ctx, w.pos_alt.grow((nn + kPosSpare) * sizeof(float4)));
    w.cur = 0;
    if (n <= 0) return LJ_OK;
    if (n == 1 || bits == 0) {
        hipLaunchKernelGGL((k_pack<KeyOp>), dim3(div_up(n, 256)), dim3(256), 0, s, x, y, z, (uint32_t)n, op,
                           w.keys[0].as<uint32_t>(), w.pos.as<float4>());
        LJ_HIP(ctx, hipGetLastError());
        return LJ_OK;
    }
    return sort_passes(ctx, x, y, z, n, op, bits, s);
}

}  // namespace lj
