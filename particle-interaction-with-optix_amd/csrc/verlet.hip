// verlet.hip -- forces over a stored neighbour list, for time stepping with list reuse.
//
// The reference rebuilds everything for every loop (it draws new particles each time,
// Cuda-version/cudaLJ.cu:602-611), so it has no counterpart.  A molecular-dynamics
// caller instead builds a list once at rc + skin (lj_pairs / lj_lbvh_pairs), reuses it
// while no particle has moved more than skin / 2 (lj_max_displacement), and evaluates
// forces over the listed pairs with the true cutoff predicate: every pair within rc
// of the current positions is then in the list, so the pair set is exactly
// lj_compute's.
#include "lj_internal.h"
#include "lj_pair.h"

namespace lj {

namespace {

// One thread per particle (row i of the CSR list, caller order).  Entries j >= n are
// ignored.  Per-block FP64 energy/virial partials; the listed and accepted pair counts
// go to res.
template <int PRED, int PREC, bool VIR>
__global__ __launch_bounds__(256) void k_list_force(const float* __restrict__ x, const float* __restrict__ y,
                                                    const float* __restrict__ z, uint32_t n,
                                                    const unsigned long long* __restrict__ rows,
                                                    const uint32_t* __restrict__ cols, PairConst c,
                                                    float* __restrict__ fx, float* __restrict__ fy,
                                                    float* __restrict__ fz, float* __restrict__ e,
                                                    double* __restrict__ partials, lj_result* __restrict__ res) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    PairSums<PREC, PRED == LJ_PRED_CUDALJ, VIR> acc;
    unsigned long long np = 0, listed = 0;
    double esum = 0.0, vsum = 0.0;
    if (i < n) {
        const float4 pi = make_float4(x[i], y[i], z[i], 0.f);
        acc.set_i(pi);
        acc.allow_exact_diff(c, pi, true);  // over the lanes with a particle
        const unsigned long long r0 = rows[i], r1 = rows[i + 1];
        for (unsigned long long q = r0; q < r1; ++q) {
            const uint32_t j = cols[q];
            if (j >= n) continue;
            ++listed;
            const float4 pj = make_float4(x[j], y[j], z[j], 0.f);
            const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const float d2 = dist2(dx, dy, dz);
            if constexpr (PRED == LJ_PRED_CUDALJ) {
                if (j == i) continue;
            }
            const int mult = pair_mult<PRED>(c, pi, pj, d2);
            if (!mult) continue;
            np += (unsigned long long)mult;
            acc.add(c, pj, dx, dy, dz, d2, mult);
        }
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
    const unsigned long long bl = block_sum_u64(listed);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = esum;
        partials[2 * blockIdx.x + 1] = vsum;
        if (bp) atomicAdd(&res->n_pairs, bp);
        if (bl) atomicAdd(&res->n_candidates, bl);
    }
}

// Largest |p - p0|^2 (FP32, a non-negative float: its bits order like the value).
__global__ __launch_bounds__(256) void k_max_disp(const float* __restrict__ x, const float* __restrict__ y,
                                                  const float* __restrict__ z, const float* __restrict__ x0,
                                                  const float* __restrict__ y0, const float* __restrict__ z0, uint32_t n,
                                                  unsigned int* __restrict__ max2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.f;
    if (i < n) {
        const float dx = x[i] - x0[i], dy = y[i] - y0[i], dz = z[i] - z0[i];
        m = dist2(dx, dy, dz);
        if (!(m >= 0.f)) m = INFINITY;  // a NaN displacement forces a rebuild
    }
#pragma unroll
    for (int d = kWaveSize / 2; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, kWaveSize));
    if (__lane_id() == 0 && m > 0.f) atomicMax(max2, __float_as_uint(m));
}

__global__ void k_zero_u32(unsigned int* p) { *p = 0u; }

}  // namespace

}  // namespace lj

using namespace lj;

extern "C" {

int lj_compute_list(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
                    const unsigned long long* row_offsets, const uint32_t* cols, float* fx, float* fy, float* fz,
                    float* e, lj_result* result, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    LJ_TRY(make_geom(ctx, p, n, &g, &c));
    if (n > 0 && (!x || !y || !z || !fx || !fy || !fz || !row_offsets || !cols))
        return fail(ctx, LJ_ERR_INVALID_ARG, "NULL pointer argument");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    lj_result* res = result;
    if (!res) {
        LJ_HIP(ctx, ctx->bin.counters.grow(sizeof(lj_result)));
        res = ctx->bin.counters.as<lj_result>();
    }
    PhaseTimer tf(ctx, LJ_PHASE_FORCE, s);
    fill_u64(&res->n_candidates, 2, 0ull, s);
    const unsigned nb = div_up(n > 0 ? n : 1, 256);
    LJ_HIP(ctx, ctx->bin.partials.grow(((size_t)nb + nb / 1024 + 1) * 2 * sizeof(double)));
    double* partials = ctx->bin.partials.as<double>();
    const bool vir = !(p->flags & LJ_FLAG_NO_VIRIAL);
    const uint32_t nn = (uint32_t)n;
#define LJ_LF(PRED, PREC, VIR)                                                                                    \
    hipLaunchKernelGGL((k_list_force<PRED, PREC, VIR>), dim3(nb), dim3(256), 0, s, x, y, z, nn, row_offsets, cols, c, \
                       fx, fy, fz, e, partials, res)
#define LJ_LF_V(PRED, PREC)            \
    do {                               \
        if (vir) LJ_LF(PRED, PREC, true); \
        else LJ_LF(PRED, PREC, false); \
    } while (0)
#define LJ_LF_P(PRED)                                          \
    switch (p->precision) {                                    \
    case LJ_PREC_MIXED: LJ_LF_V(PRED, LJ_PREC_MIXED); break;   \
    case LJ_PREC_FP32: LJ_LF_V(PRED, LJ_PREC_FP32); break;     \
    default: LJ_LF_V(PRED, LJ_PREC_FP32_COMP); break;          \
    }
    switch (p->predicate) {
    case LJ_PRED_AABB: LJ_LF_P(LJ_PRED_AABB); break;
    case LJ_PRED_CUSTOM: LJ_LF_P(LJ_PRED_CUSTOM); break;
    case LJ_PRED_CUDALJ: LJ_LF_P(LJ_PRED_CUDALJ); break;
    case LJ_PRED_SPHERE: LJ_LF_P(LJ_PRED_SPHERE); break;
    case LJ_PRED_TRIANGLE: LJ_LF_P(LJ_PRED_TRIANGLE); break;
    }
#undef LJ_LF_P
#undef LJ_LF_V
#undef LJ_LF
    finalize_energy(partials, nb, res, s);
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

int lj_max_displacement(lj_context* ctx, const float* x, const float* y, const float* z, const float* x0,
                        const float* y0, const float* z0, int64_t n, float* max2, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    if (n < 0 || n > 0x7fffffffll) return fail(ctx, LJ_ERR_INVALID_ARG, "particle count out of range");
    if (!max2 || (n > 0 && (!x || !y || !z || !x0 || !y0 || !z0)))
        return fail(ctx, LJ_ERR_INVALID_ARG, "NULL pointer argument");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    unsigned int* m = reinterpret_cast<unsigned int*>(max2);
    // a fill kernel, not hipMemsetAsync: the call may be captured in a graph
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, m);
    if (n > 0)
        hipLaunchKernelGGL(k_max_disp, dim3(div_up(n, 256)), dim3(256), 0, s, x, y, z, x0, y0, z0, (uint32_t)n, m);
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

}  // extern "C"
