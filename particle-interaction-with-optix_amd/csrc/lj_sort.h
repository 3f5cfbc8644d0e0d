// lj_sort.h -- stable LSD radix sort of particles by a 32-bit key, carrying the payload
// (x, y, z, original index) through every pass.
//
// One pass over 8 key bits is two kernels around a device-wide scan:
//   histogram  per 2048-key tile, digit-major table hist[d * ntiles + tile]
//   scan       exclusive scan of the table: hist[d, tile] = first output slot
//   scatter    each tile ranks its keys stably (input order inside a digit), stages
//              key + payload in LDS in digit order, and writes every digit's run to
//              consecutive output slots, so stores are contiguous runs rather than
//              one scattered 20-byte record per lane.
// The first pass computes the keys from the positions with a key functor (cell key,
// batched cell key, Morton code) and reads x, y, z in input order, so no pass ever
// gathers positions through a permutation.
#pragma once

#include "lj_internal.h"

namespace lj {

namespace sortk {

constexpr int kWave = 64;
constexpr int kBits = 8;
constexpr int kRadix = 1 << kBits;
constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;  // 2048 keys per tile
constexpr int kWaves = kThreads / kWave;

__device__ inline unsigned rank_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Key + payload of element idx, from positions (first pass) or from the previous pass.
template <bool FIRST, class KeyOp>
__device__ inline void load_elem(uint32_t idx, const float* __restrict__ x, const float* __restrict__ y,
                                 const float* __restrict__ z, const uint32_t* __restrict__ kin,
                                 const float4* __restrict__ pin, const KeyOp& op, uint32_t& key, float4& pay) {
    if constexpr (FIRST) {
        const float px = x[idx], py = y[idx], pz = z[idx];
        key = op(px, py, pz, idx);
        pay = make_float4(px, py, pz, __uint_as_float(idx));
    } else {
        key = kin[idx];
        pay = pin[idx];
    }
}

template <bool FIRST, class KeyOp>
__global__ __launch_bounds__(kThreads) void k_hist(const float* __restrict__ x, const float* __restrict__ y,
                                                   const float* __restrict__ z, const uint32_t* __restrict__ kin,
                                                   uint32_t n, KeyOp op, int shift, uint32_t ntiles,
                                                   uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t idx = base + r * kThreads + threadIdx.x;
        if (idx < n) {
            uint32_t key;
            if constexpr (FIRST) key = op(x[idx], y[idx], z[idx], idx);
            else key = kin[idx];
            atomicAdd(&h[(key >> shift) & (kRadix - 1)], 1u);
        }
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

template <bool FIRST, class KeyOp>
__global__ __launch_bounds__(kThreads) void k_scatter(const float* __restrict__ x, const float* __restrict__ y,
                                                      const float* __restrict__ z, const uint32_t* __restrict__ kin,
                                                      const float4* __restrict__ pin, uint32_t n, KeyOp op, int shift,
                                                      uint32_t ntiles, const uint32_t* __restrict__ hist,
                                                      uint32_t* __restrict__ kout, float4* __restrict__ pout) {
    __shared__ uint32_t skey[kTile];
    __shared__ float4 spay[kTile];
    __shared__ uint32_t dstart[kRadix];  // first staging slot of each digit in this tile
    __shared__ uint32_t gbase[kRadix];   // first output slot of each digit for this tile
    __shared__ uint32_t run[kRadix];     // keys of each digit already ranked in earlier rounds
    __shared__ uint32_t wcnt[kWaves][kRadix];
    __shared__ uint32_t wsum[kWaves];
    const unsigned tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    const uint32_t base = blockIdx.x * kTile;
    const uint32_t count = n - base < (uint32_t)kTile ? n - base : (uint32_t)kTile;

    uint32_t key[kItems];
    float4 pay[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t j = r * kThreads + tid;
        if (j < count) load_elem<FIRST>(base + j, x, y, z, kin, pin, op, key[r], pay[r]);
        else key[r] = 0;
    }
    // tile digit counts -> exclusive starts (thread d owns digit d)
    run[tid] = 0;
    gbase[tid] = hist[(uint64_t)tid * ntiles + blockIdx.x];
    dstart[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kItems; ++r)
        if (r * kThreads + tid < count) atomicAdd(&dstart[(key[r] >> shift) & (kRadix - 1)], 1u);
    __syncthreads();
    {
        const uint32_t v = dstart[tid];
        uint32_t incl = v;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d, kWave);
            if (lane >= (unsigned)d) incl += o;
        }
        if (lane == kWave - 1) wsum[wave] = incl;
        __syncthreads();
        uint32_t off = 0;
        for (unsigned q = 0; q < wave; ++q) off += wsum[q];
        dstart[tid] = off + incl - v;
    }
    // stable ranks, round by round; (round, wave, lane) is the input order
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        const bool valid = r * kThreads + tid < count;
        const uint32_t d = (key[r] >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const unsigned rk = rank_below(peers);
        if (valid && rk == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t acc = run[tid];
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = wcnt[w][tid];
                wcnt[w][tid] = acc;
                acc += c;
            }
            run[tid] = acc;
        }
        __syncthreads();
        if (valid) {
            const uint32_t slot = dstart[d] + wcnt[wave][d] + rk;
            skey[slot] = key[r];
            spay[slot] = pay[r];
        }
        __syncthreads();
    }
    // write the staged tile: consecutive slots of one digit go to consecutive outputs
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t p = r * kThreads + tid;
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

// Sorts n particles by op's key (the low `bits` bits).  The sorted keys land in
// ctx->bin.keys[ctx->bin.cur] and the payload (x, y, z, bits of the original index)
// in ctx->bin.pos, which has room for n + 1 entries (entry n is not written).
template <class KeyOp>
int sort_particles(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const KeyOp& op,
                   int bits, hipStream_t s) {
    using namespace sortk;
    BinWork& w = ctx->bin;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    LJ_HIP(ctx, w.keys[0].grow(nn * 4));
    LJ_HIP(ctx, w.keys[1].grow(nn * 4));
    // one slot beyond the last particle: the force loops read pairs of slots (t, t+1)
    // without clamping t+1 and use the second only when it is in range
    LJ_HIP(ctx, w.pos.grow((nn + 1) * sizeof(float4)));
    LJ_HIP(ctx, w.pos_alt.grow((nn + 1) * sizeof(float4)));
    w.cur = 0;
    if (n <= 0) return LJ_OK;
    const int passes = (n > 1) ? (bits + kBits - 1) / kBits : 0;
    if (passes == 0) {
        hipLaunchKernelGGL((k_pack<KeyOp>), dim3(div_up(n, 256)), dim3(256), 0, s, x, y, z, (uint32_t)n, op,
                           w.keys[0].as<uint32_t>(), w.pos.as<float4>());
        LJ_HIP(ctx, hipGetLastError());
        return LJ_OK;
    }
    const unsigned ntiles = div_up(n, kTile);
    LJ_HIP(ctx, w.hist.grow((size_t)kRadix * ntiles * 4));
    uint32_t* hist = w.hist.as<uint32_t>();
    // the last pass must write ctx->bin.pos
    float4* pbuf[2] = {w.pos.as<float4>(), w.pos_alt.as<float4>()};
    int kcur = 0;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kBits;
        const int kdst = p == 0 ? 0 : kcur ^ 1;
        float4* pdst = pbuf[(passes - 1 - p) & 1];
        const float4* psrc = pbuf[(passes - p) & 1];
        if (p == 0) {
            hipLaunchKernelGGL((k_hist<true, KeyOp>), dim3(ntiles), dim3(kThreads), 0, s, x, y, z,
                               (const uint32_t*)nullptr, (uint32_t)n, op, shift, ntiles, hist);
        } else {
            hipLaunchKernelGGL((k_hist<false, KeyOp>), dim3(ntiles), dim3(kThreads), 0, s, x, y, z,
                               (const uint32_t*)w.keys[kcur].as<uint32_t>(), (uint32_t)n, op, shift, ntiles, hist);
        }
        LJ_TRY(scan_u32(ctx, hist, hist, (int64_t)kRadix * ntiles, nullptr, s));
        if (p == 0) {
            // pass 0 reads positions; it writes keys[1] so that keys[0] stays free
            hipLaunchKernelGGL((k_scatter<true, KeyOp>), dim3(ntiles), dim3(kThreads), 0, s, x, y, z,
                               (const uint32_t*)nullptr, (const float4*)nullptr, (uint32_t)n, op, shift, ntiles,
                               (const uint32_t*)hist, w.keys[1].as<uint32_t>(), pdst);
            kcur = 1;
        } else {
            hipLaunchKernelGGL((k_scatter<false, KeyOp>), dim3(ntiles), dim3(kThreads), 0, s, x, y, z,
                               (const uint32_t*)w.keys[kcur].as<uint32_t>(), psrc, (uint32_t)n, op, shift, ntiles,
                               (const uint32_t*)hist, w.keys[kdst].as<uint32_t>(), pdst);
            kcur = kdst;
        }
    }
    w.cur = kcur;
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

}  // namespace lj
