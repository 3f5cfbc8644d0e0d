// force.hip -- cell-grid Lennard-Jones forces/energies, neighbour lists and reorder.
//
// Reference: ComputeParticleInterationsParPartNoLoop (Cuda-version/cudaLJ.cu:299-341)
// computes, per particle in cell order, the potential against every particle of
// the clipped 27-cell stencil, with no cutoff and sigma = epsilon = 1, and no force.
// Here the pair set is chosen by the predicate of lj_params (the canonical one is
// the OptiX AABB test, SDK/optixMultiSphereAABB/optixLJ.cu:162), the force is the
// analytic gradient of the same potential, and per-particle outputs are written
// in the caller's order.
//
// Cell order is z-major (cudaLJ.cu:127-134), so the three x-cells of one (y,z) row
// of the stencil are one contiguous slot range: the 27 cells are 9 ranges.
#include "lj_internal.h"
#include "lj_pair.h"

namespace lj {

namespace {

__device__ inline void cell_coords(const GridGeom& g, uint32_t key, int32_t& cx, int32_t& cy, int32_t& cz) {
    cx = (int32_t)(key % (uint32_t)g.g[0]);
    const uint32_t r = key / (uint32_t)g.g[0];
    cy = (int32_t)(r % (uint32_t)g.g[1]);
    cz = (int32_t)(r / (uint32_t)g.g[1]);
}

// One thread per sorted slot k.  The 27 cells are visited as 9 contiguous x-rows in
// the z -> y order of cudaLJ.cu:319-321.
template <int PRED, int PREC, bool VIR>
__global__ __launch_bounds__(256) void k_force_simple(const float4* __restrict__ pos, const uint32_t* __restrict__ cell_of,
                                                      const uint32_t* __restrict__ off, uint32_t n, GridGeom g,
                                                      PairConst c, float* __restrict__ fx, float* __restrict__ fy,
                                                      float* __restrict__ fz, float* __restrict__ e,
                                                      double* __restrict__ partials,
                                                      unsigned long long* __restrict__ npairs) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    PairSums<PREC, PRED == LJ_PRED_CUDALJ, VIR> acc;
    unsigned long long np = 0;
    double esum = 0.0, vsum = 0.0;
    if (k < n) {
        const float4 pi = pos[k];
        acc.set_i(pi);
        int32_t cx, cy, cz;
        cell_coords(g, cell_of[k], cx, cy, cz);
        const int32_t x0 = cx > 0 ? cx - 1 : 0, x1 = cx < g.g[0] - 1 ? cx + 1 : g.g[0] - 1;
        for (int32_t zz = cz - 1; zz <= cz + 1; ++zz) {
            if (zz < 0 || zz >= g.g[2]) continue;
            for (int32_t yy = cy - 1; yy <= cy + 1; ++yy) {
                if (yy < 0 || yy >= g.g[1]) continue;
                const uint32_t row = (uint32_t)((zz * g.g[1] + yy) * g.g[0]);
                const uint32_t b = off[row + x0], en = off[row + x1 + 1];
                for (uint32_t t = b; t < en; ++t) {
                    const float4 pj = pos[t];
                    const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                    const float d2 = dist2(dx, dy, dz);
                    if (t == k || !accept<PRED>(c, d2)) continue;
                    ++np;
                    acc.add(c, pj, dx, dy, dz, d2);
                }
            }
        }
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
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = esum;
        partials[2 * blockIdx.x + 1] = vsum;
        if (npairs && bp) atomicAdd(npairs, bp);
    }
}

// Final FP64 reduction of the per-block partials, in a fixed order.
__global__ __launch_bounds__(256) void k_finalize(const double* __restrict__ partials, uint32_t nblocks,
                                                  lj_result* __restrict__ res) {
    double a = 0.0, b = 0.0;
    for (uint32_t q = threadIdx.x; q < nblocks; q += 256) {
        a += partials[2 * q];
        b += partials[2 * q + 1];
    }
    block_sum2(a, b);
    if (threadIdx.x == 0) {
        res->e_total = 0.5 * a;
        res->virial = 0.5 * b;
    }
}

// ------------------------------------------------------------------ pair lists
// Pass 1: number of accepted partners of each particle, stored by original index.
template <int PRED>
__global__ __launch_bounds__(256) void k_pair_count(const float4* __restrict__ pos, const uint32_t* __restrict__ cell_of,
                                                    const uint32_t* __restrict__ off, uint32_t n, GridGeom g,
                                                    PairConst c, uint32_t* __restrict__ rowcnt) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 pi = pos[k];
    int32_t cx, cy, cz;
    cell_coords(g, cell_of[k], cx, cy, cz);
    const int32_t x0 = cx > 0 ? cx - 1 : 0, x1 = cx < g.g[0] - 1 ? cx + 1 : g.g[0] - 1;
    uint32_t cnt = 0;
    for (int32_t zz = cz - 1; zz <= cz + 1; ++zz) {
        if (zz < 0 || zz >= g.g[2]) continue;
        for (int32_t yy = cy - 1; yy <= cy + 1; ++yy) {
            if (yy < 0 || yy >= g.g[1]) continue;
            const uint32_t row = (uint32_t)((zz * g.g[1] + yy) * g.g[0]);
            for (uint32_t t = off[row + x0]; t < off[row + x1 + 1]; ++t) {
                const float4 pj = pos[t];
                const float d2 = dist2(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
                if (t != k && accept<PRED>(c, d2)) ++cnt;
            }
        }
    }
    rowcnt[__float_as_uint(pi.w)] = cnt;
}

// Pass 2: the partners' original indices, then each row sorted ascending.
template <int PRED>
__global__ __launch_bounds__(256) void k_pair_fill(const float4* __restrict__ pos, const uint32_t* __restrict__ cell_of,
                                                   const uint32_t* __restrict__ off, uint32_t n, GridGeom g,
                                                   PairConst c, const unsigned long long* __restrict__ rows,
                                                   uint32_t* __restrict__ cols) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 pi = pos[k];
    const uint32_t i = __float_as_uint(pi.w);
    int32_t cx, cy, cz;
    cell_coords(g, cell_of[k], cx, cy, cz);
    const int32_t x0 = cx > 0 ? cx - 1 : 0, x1 = cx < g.g[0] - 1 ? cx + 1 : g.g[0] - 1;
    const unsigned long long r0 = rows[i];
    unsigned long long w = r0;
    for (int32_t zz = cz - 1; zz <= cz + 1; ++zz) {
        if (zz < 0 || zz >= g.g[2]) continue;
        for (int32_t yy = cy - 1; yy <= cy + 1; ++yy) {
            if (yy < 0 || yy >= g.g[1]) continue;
            const uint32_t row = (uint32_t)((zz * g.g[1] + yy) * g.g[0]);
            for (uint32_t t = off[row + x0]; t < off[row + x1 + 1]; ++t) {
                const float4 pj = pos[t];
                const float d2 = dist2(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
                if (t != k && accept<PRED>(c, d2)) cols[w++] = __float_as_uint(pj.w);
            }
        }
    }
    // insertion sort of the row (rows hold tens of entries)
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

__global__ __launch_bounds__(256) void k_unpack(const float4* __restrict__ pos, uint32_t n, float* __restrict__ xo,
                                                float* __restrict__ yo, float* __restrict__ zo) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 p = pos[k];
    xo[k] = p.x;
    yo[k] = p.y;
    zo[k] = p.z;
}

template <int PRED>
void launch_force(int prec, bool vir, dim3 grid, hipStream_t s, const float4* pos, const uint32_t* cell_of,
                  const uint32_t* off, uint32_t n, const GridGeom& g, const PairConst& c, float* fx, float* fy,
                  float* fz, float* e, double* partials, unsigned long long* npairs) {
#define LJ_LAUNCH(PREC, VIR)                                                                                  \
    hipLaunchKernelGGL((k_force_simple<PRED, PREC, VIR>), grid, dim3(256), 0, s, pos, cell_of, off, n, g, c, fx, \
                       fy, fz, e, partials, npairs)
    switch (prec) {
    case LJ_PREC_MIXED: if (vir) LJ_LAUNCH(LJ_PREC_MIXED, true); else LJ_LAUNCH(LJ_PREC_MIXED, false); break;
    case LJ_PREC_FP32: if (vir) LJ_LAUNCH(LJ_PREC_FP32, true); else LJ_LAUNCH(LJ_PREC_FP32, false); break;
    default: if (vir) LJ_LAUNCH(LJ_PREC_FP32_COMP, true); else LJ_LAUNCH(LJ_PREC_FP32_COMP, false); break;
    }
#undef LJ_LAUNCH
}

}  // namespace

void finalize_energy(const double* partials, unsigned nblocks, lj_result* res, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, partials, (uint32_t)nblocks, res);
}

// Scalar results live in the caller's lj_result or, when it is NULL, in a scratch one.
static int result_slot(lj_context* ctx, lj_result* user, lj_result** out) {
    if (user) {
        *out = user;
        return LJ_OK;
    }
    LJ_HIP(ctx, ctx->bin.counters.grow(sizeof(lj_result)));
    *out = ctx->bin.counters.as<lj_result>();
    return LJ_OK;
}

int compute_cellgrid(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
                     const GridGeom& g, const PairConst& c, float* fx, float* fy, float* fz, float* e,
                     lj_result* result, hipStream_t s) {
    {
        PhaseTimer t(ctx, LJ_PHASE_BIN, s);
        LJ_TRY(bin_particles(ctx, x, y, z, n, g, s));
    }
    lj_result* res;
    LJ_TRY(result_slot(ctx, result, &res));
    {
        PhaseTimer t(ctx, LJ_PHASE_OTHER, s);
        LJ_TRY(count_candidates(ctx, g, &res->n_candidates, s));
        fill_u64(&res->n_pairs, 1, 0ull, s);
    }
    const unsigned nb = div_up(n > 0 ? n : 1, 256);
    LJ_HIP(ctx, ctx->bin.partials.grow((size_t)nb * 2 * sizeof(double)));
    double* partials = ctx->bin.partials.as<double>();
    const int prec = p->precision;
    const bool vir = !(p->flags & LJ_FLAG_NO_VIRIAL);
    const float4* pos = ctx->bin.pos.as<float4>();
    const uint32_t* cell_of = ctx->bin.sorted_keys();
    const uint32_t* off = ctx->bin.offsets.as<uint32_t>();
    const uint32_t nn = (uint32_t)n;
    {
    PhaseTimer tf(ctx, LJ_PHASE_FORCE, s);
    switch (p->predicate) {
    case LJ_PRED_AABB:
        launch_force<LJ_PRED_AABB>(prec, vir, dim3(nb), s, pos, cell_of, off, nn, g, c, fx, fy, fz, e, partials, &res->n_pairs);
        break;
    case LJ_PRED_CUSTOM:
        launch_force<LJ_PRED_CUSTOM>(prec, vir, dim3(nb), s, pos, cell_of, off, nn, g, c, fx, fy, fz, e, partials, &res->n_pairs);
        break;
    case LJ_PRED_CUDALJ:
        launch_force<LJ_PRED_CUDALJ>(prec, vir, dim3(nb), s, pos, cell_of, off, nn, g, c, fx, fy, fz, e, partials, &res->n_pairs);
        break;
    case LJ_PRED_SPHERE:
        launch_force<LJ_PRED_SPHERE>(prec, vir, dim3(nb), s, pos, cell_of, off, nn, g, c, fx, fy, fz, e, partials, &res->n_pairs);
        break;
    case LJ_PRED_TRIANGLE:
        launch_force<LJ_PRED_TRIANGLE>(prec, vir, dim3(nb), s, pos, cell_of, off, nn, g, c, fx, fy, fz, e, partials, &res->n_pairs);
        break;
    }
    }
    PhaseTimer to(ctx, LJ_PHASE_OTHER, s);
    finalize_energy(partials, nb, res, s);
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

int pairs_cellgrid(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
                   const GridGeom& g, const PairConst& c, unsigned long long* rows, uint32_t* cols,
                   unsigned long long capacity, unsigned long long* needed, hipStream_t s) {
    LJ_TRY(bin_particles(ctx, x, y, z, n, g, s));
    const unsigned nb = div_up(n > 0 ? n : 1, 256);
    const float4* pos = ctx->bin.pos.as<float4>();
    const uint32_t* cell_of = ctx->bin.sorted_keys();
    const uint32_t* off = ctx->bin.offsets.as<uint32_t>();
    const uint32_t nn = (uint32_t)n;
    // row counts (u32, by original index) -> u64 offsets
    LJ_HIP(ctx, ctx->bin.rows.grow(((size_t)n + 1) * sizeof(uint32_t)));
    uint32_t* rowcnt = ctx->bin.rows.as<uint32_t>();
    switch (p->predicate) {
#define LJ_COUNT(P) case P: hipLaunchKernelGGL(k_pair_count<P>, dim3(nb), dim3(256), 0, s, pos, cell_of, off, nn, g, c, rowcnt); break;
        LJ_COUNT(LJ_PRED_AABB) LJ_COUNT(LJ_PRED_CUSTOM) LJ_COUNT(LJ_PRED_CUDALJ) LJ_COUNT(LJ_PRED_SPHERE) LJ_COUNT(LJ_PRED_TRIANGLE)
#undef LJ_COUNT
    }
    LJ_TRY(scan_u32_to_u64(ctx, rowcnt, rows, n, s));
    unsigned long long total = 0;
    LJ_HIP(ctx, hipMemcpyAsync(&total, rows + n, sizeof total, hipMemcpyDeviceToHost, s));
    LJ_HIP(ctx, hipStreamSynchronize(s));
    if (needed) *needed = total;
    if (total > capacity)
        return fail(ctx, LJ_ERR_CAPACITY, "pair list needs %llu entries, capacity is %llu", total, capacity);
    if (total == 0) return LJ_OK;
    switch (p->predicate) {
#define LJ_FILL(P) case P: hipLaunchKernelGGL(k_pair_fill<P>, dim3(nb), dim3(256), 0, s, pos, cell_of, off, nn, g, c, (const unsigned long long*)rows, cols); break;
        LJ_FILL(LJ_PRED_AABB) LJ_FILL(LJ_PRED_CUSTOM) LJ_FILL(LJ_PRED_CUDALJ) LJ_FILL(LJ_PRED_SPHERE) LJ_FILL(LJ_PRED_TRIANGLE)
#undef LJ_FILL
    }
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

int unpack_sorted(lj_context* ctx, int64_t n, float* xo, float* yo, float* zo, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_unpack, dim3(div_up(n, 256)), dim3(256), 0, s, ctx->bin.pos.as<float4>(), (uint32_t)n, xo,
                           yo, zo);
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}

}  // namespace lj

using namespace lj;

extern "C" {

int lj_compute(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
               float* fx, float* fy, float* fz, float* e, lj_result* result, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    LJ_TRY(make_geom(ctx, p, n, &g, &c));
    if (n > 0 && (!x || !y || !z || !fx || !fy || !fz))
        return fail(ctx, LJ_ERR_INVALID_ARG, "positions and forces must not be NULL");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    return compute_cellgrid(ctx, x, y, z, n, p, g, c, fx, fy, fz, e, result, (hipStream_t)stream);
}

int lj_pairs(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
             unsigned long long* row_offsets, uint32_t* cols, unsigned long long capacity,
             unsigned long long* needed, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    LJ_TRY(make_geom(ctx, p, n, &g, &c));
    if (!row_offsets || (n > 0 && (!x || !y || !z)) || (capacity > 0 && !cols))
        return fail(ctx, LJ_ERR_INVALID_ARG, "NULL pointer argument");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    return pairs_cellgrid(ctx, x, y, z, n, p, g, c, row_offsets, cols, capacity, needed, (hipStream_t)stream);
}

int lj_reorder(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
               float* xo, float* yo, float* zo, uint32_t* perm, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    LJ_TRY(make_geom(ctx, p, n, &g, &c));
    if (n > 0 && (!x || !y || !z || !xo || !yo || !zo))
        return fail(ctx, LJ_ERR_INVALID_ARG, "NULL pointer argument");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    LJ_TRY(bin_particles(ctx, x, y, z, n, g, s));
    LJ_TRY(unpack_sorted(ctx, n, xo, yo, zo, s));
    if (perm && n > 0)
        LJ_HIP(ctx, hipMemcpyAsync(perm, ctx->bin.perm(), (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    return LJ_OK;
}

}  // extern "C"
