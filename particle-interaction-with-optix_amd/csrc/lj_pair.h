// lj_pair.h -- per-pair Lennard-Jones arithmetic and block reductions shared by the
// cell-grid, LBVH and batched kernels (one definition, so every path computes the
// same terms for the same pair).
#pragma once

#include "lj_device.h"

namespace lj {

constexpr int kWaveSize = 64;

// Per-particle accumulators of one precision mode (LJ_PREC_*).
//   MIXED      FP64 sums of FP64 pair terms
//   FP32       plain FP32 sums
//   FP32_COMP  FP32 sums with the TwoSum error term kept apart (Knuth, TAOCP 2,
//              4.2.2; Ogita, Rump, Oishi, "Accurate sum and dot product", 2005)
template <int PREC>
struct Acc;

template <>
struct Acc<LJ_PREC_MIXED> {
    double s = 0.0;
    __device__ inline void add(double v) { s += v; }
    __device__ inline void fma(double a, double b) { s = __builtin_fma(a, b, s); }
    __device__ inline double value() const { return s; }
};

template <>
struct Acc<LJ_PREC_FP32> {
    float s = 0.f;
    __device__ inline void add(float v) { s += v; }
    __device__ inline void fma(float a, float b) { s = __builtin_fmaf(a, b, s); }
    __device__ inline float value() const { return s; }
};

template <>
struct Acc<LJ_PREC_FP32_COMP> {
    float s = 0.f, c = 0.f;
    __device__ inline void add(float v) {
        const float t = s + v;
        const float bp = t - s;
        c += (s - (t - bp)) + (v - bp);
        s = t;
    }
    __device__ inline void fma(float a, float b) { add(a * b); }
    __device__ inline float value() const { return s + c; }
};

// FP64 reciprocal: hardware estimate plus one Newton-Raphson step.  Measured on
// MI355X over 4M log-uniform inputs in [1e-20, 1e4]: the estimate is within 2^-24.4
// relative, one step within 2.3e-15 (a second step would make it correctly rounded,
// which the FP32 outputs cannot show).
__device__ inline double rcp_f64(double a) {
    const double y = __builtin_amdgcn_rcp(a);
    const double e = __builtin_fma(-a, y, 1.0);
    return __builtin_fma(y, e, y);
}

// Accumulates one accepted pair d = p_j - p_i into the per-particle sums:
//   t = s6 (2 s6 - 1),  w = t / r2   (F_i = -24 eps sum w d),
//   energy term s6 (s6 - 1)          (e_i = 4 eps sum),
//   virial term t                    (sum_j w r2 = sum_j t).
// s6 = (sigma^2 / r2)^3 with r2 = |d|^2 (+ softening in CUDALJ mode).  In MIXED
// precision d is taken in FP64 from the FP32 positions, where the difference of two
// floats is exact; in FP32 modes d is the FP32 difference used by the predicate.
template <int PREC, bool SOFT, bool VIR>
struct PairSums {
    Acc<PREC> ax, ay, az, ae, av;
    double px = 0.0, py = 0.0, pz = 0.0;  // p_i in FP64 (MIXED)
    bool exact = false;                   // wave-uniform, see allow_exact_diff
    __device__ inline void set_i(float4 pi) {
        px = pi.x;
        py = pi.y;
        pz = pi.z;
    }
    // Accepted partners of a cutoff predicate lie within rc of p_i on every axis.  If
    // |p_i| >= 2 rc on every axis (so p_j has the sign of p_i and lies in [p_i/2, 2 p_i]),
    // Sterbenz's lemma makes the FP32 difference p_j - p_i exact, and converting it
    // gives the FP64 difference bit for bit with one instruction per axis instead of
    // two.  The fast form is used only when it holds for every calling lane, so the
    // choice is a uniform branch: call it from every lane of the wave (idle lanes with
    // active = false), or from the lanes of a branch that alone uses the flag.
    __device__ inline void allow_exact_diff(const PairConst& c, float4 pi, bool active) {
        const float reach = 2.001f * c.rc;
        const bool ok = !active || (!SOFT && fabsf(pi.x) >= reach && fabsf(pi.y) >= reach && fabsf(pi.z) >= reach);
        exact = __all(ok);
    }
    // mult: how many times the pair counts (the TRIANGLE ray model can count a pair
    // more than once); 1 for every other predicate.
    __device__ inline void add(const PairConst& c, float4 pj, float dx, float dy, float dz, float d2, int mult = 1) {
        if constexpr (PREC == LJ_PREC_MIXED) {
            double ddx, ddy, ddz;
            if (exact) {
                ddx = dx;
                ddy = dy;
                ddz = dz;
            } else {
                ddx = (double)pj.x - px;
                ddy = (double)pj.y - py;
                ddz = (double)pj.z - pz;
            }
            double r2 = __builtin_fma(ddz, ddz, __builtin_fma(ddy, ddy, ddx * ddx));
            if constexpr (SOFT) r2 += c.softd;
            const double ir2 = rcp_f64(r2);
            const double s2 = c.sig2d * ir2;
            const double s6 = s2 * s2 * s2;
            double u = __builtin_fma(s6, s6, -s6);  // energy term s6 (s6 - 1)
            double t = __builtin_fma(s6, s6, u);    // s6 (2 s6 - 1)
            if (mult != 1) {
                u *= (double)mult;
                t *= (double)mult;
            }
            const double w = t * ir2;
            ax.fma(w, ddx);
            ay.fma(w, ddy);
            az.fma(w, ddz);
            ae.add(u);
            if constexpr (VIR) av.add(t);
        } else {
            float r2 = d2;
            if constexpr (SOFT) r2 += c.soft;
            const float ir2 = __builtin_amdgcn_rcpf(r2);
            const float s2 = c.sig2 * ir2;
            const float s6 = s2 * s2 * s2;
            float t = __builtin_fmaf(s6 + s6, s6, -s6);
            float u = __builtin_fmaf(s6, s6, -s6);
            if (mult != 1) {
                t *= (float)mult;
                u *= (float)mult;
            }
            const float w = t * ir2;
            ax.fma(w, dx);
            ay.fma(w, dy);
            az.fma(w, dz);
            ae.add(u);
            if constexpr (VIR) av.add(t);
        }
    }
};

// Store of one per-particle output at its caller index (or, for lj_compute_sorted, at
// its sorted slot, where the stores are consecutive).  Sorted slots map to random
// caller indices, so every caller-order store is a partial write of its own line.  When the
// outputs fit in the L2s, ordinary stores let those lines fill up there before they go
// to memory; when they do not, a non-temporal store keeps each line out of the caches,
// where it would displace the neighbour positions the force loops re-read.  Measured
// (force kernels, ordinary -> non-temporal): C2 1M particles 0.145 -> 0.161 ms, C5 4.2M
// 0.91 -> 0.88 ms (LBVH 2.65 -> 2.60), C3 16.7M 4.26 -> 4.03 ms, and the batched C4,
// whose systems each write a range of at most 256 KB, 8.1 -> 9.1 ms.  Beyond the L2s the
// caller-order outputs are now staged (store_staged, below); non-temporal scalar stores
// remain for lj_compute_sorted's consecutive ones.  The kind is a template parameter of
// the kernels (chosen on the host by non_temporal_outputs): with a run-time branch
// between the two forms the compiler merges them into one ordinary store.
template <bool NT>
__device__ inline void store_out(float* p, float v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Whether a launch writing n particles' outputs (four floats each, at random indices)
// should store them non-temporally: when they exceed the chip's L2 (8 XCDs x 4 MiB).
constexpr bool non_temporal_outputs(int64_t n) { return (uint64_t)(n > 0 ? n : 0) * 16u > (32ull << 20); }

// Staged outputs.  Beyond the L2s, a particle's four outputs at a shuffled index are
// four partial writes of four sectors; written instead as one 16-byte record (fx, fy,
// fz, e) at that index of a staging array, they are one partial write, and
// split_outputs then copies the staging array into the four caller arrays with
// consecutive loads and stores.  C3 (16.7M particles, outputs in caller order): force
// kernel 4.07 -> 3.81 ms with the split pass included, step 4.83 -> 4.61 ms.
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ inline void store_staged(float4* p, float a, float b, float c, float d) {
    const f32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<f32x4*>(p));
}

template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int d = kWaveSize / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWaveSize);
    return v;
}

// Blocks of up to 256 threads (a whole number of waves): deterministic tree sum of two
// doubles; thread 0 holds the result.  Every thread must call it; the closing barrier
// lets a block call it again (its LDS slots are reused).
__device__ inline void block_sum2(double& a, double& b) {
    __shared__ double sh[2][256 / kWaveSize];
    a = wave_sum(a);
    b = wave_sum(b);
    const unsigned w = threadIdx.x / kWaveSize;
    if (__lane_id() == 0) {
        sh[0][w] = a;
        sh[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = 0.0;
        b = 0.0;
        for (unsigned q = 0; q < blockDim.x / kWaveSize; ++q) {
            a += sh[0][q];
            b += sh[1][q];
        }
    }
    __syncthreads();
}

// Sum over the block (thread 0 holds it), callable repeatedly like block_sum2.
__device__ inline unsigned long long block_sum_u64(unsigned long long v) {
    __shared__ unsigned long long sh[256 / kWaveSize];
    v = wave_sum(v);
    if (__lane_id() == 0) sh[threadIdx.x / kWaveSize] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (unsigned q = 0; q < blockDim.x / kWaveSize; ++q) t += sh[q];
    __syncthreads();
    return t;
}

}  // namespace lj
