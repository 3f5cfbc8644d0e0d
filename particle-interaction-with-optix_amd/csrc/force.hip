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
// Cell order is z-major (cudaLJ.cu:127-134), so the x-cells of one (y,z) row of the
// stencil are one contiguous slot range: the 27 cells are 9 ranges (the stencil loop
// itself is stencil_pairs in lj_stencil.h).
#include <cmath>

#include "lj_stencil.h"

namespace lj {

namespace {

// One thread per sorted slot k, blocks of `block` threads (<= 256), the lane queues in
// dynamic LDS (depth x block entries of 16 bytes, one block of rings per wave; none
// without QUEUE).  Outputs go to the caller's order (pi.w holds the original index),
// or, when `so` is set (lj_compute_sorted), to the sorted slot k together with the
// particle's position and original index; per-block FP64 energy and virial partials are
// reduced later in a fixed order.
// OUT: how the outputs are stored -- 0 ordinary stores, 1 non-temporal stores, 2 one
// non-temporal staged record per particle at its caller index (large outputs at
// shuffled indices; never with `so`).
template <int PRED, int PREC, bool VIR, bool QUEUE, int OUT>
__global__ __launch_bounds__(kMaxBlock, 8) void k_force(const float4* __restrict__ pos, const uint32_t* __restrict__ cell_of,
                                                     const uint32_t* __restrict__ off, uint32_t n, GridGeom g, int Rx,
                                                     int R, Trim tr, PairConst c, uint32_t depth, float* __restrict__ fx,
                                                     float* __restrict__ fy, float* __restrict__ fz,
                                                     float* __restrict__ e, double* __restrict__ partials,
                                                     unsigned long long* __restrict__ npairs, SortedOut so,
                                                     float4* __restrict__ stage) {
    extern __shared__ QEntry queue[];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t k = blk * blockDim.x + threadIdx.x;
    // the slot rebuilt after the pair loop from wave-uniform values and the lane id, so
    // that no per-lane index stays live across it (64 VGPRs are the budget)
    const uint32_t wbase = blk * blockDim.x + (__builtin_amdgcn_readfirstlane(threadIdx.x) & ~(kWaveSize - 1u));
    const bool active = k < n;
    PairSums<PREC, PRED == LJ_PRED_CUDALJ, VIR> acc;
    const float4 pi = active ? pos[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    acc.set_i(pi);
    acc.allow_exact_diff(c, pi, active);
    const uint32_t np = stencil_pairs<PRED, PREC, VIR, QUEUE>(pos, off, k, active, pi, g, 0u, active ? cell_of[k] : 0u,
                                                              Rx, R, tr, c,
                                                              queue + (threadIdx.x / kWaveSize) * depth * kWaveSize +
                                                                  __lane_id(),
                                                              depth, acc);
    double esum = 0.0, vsum = 0.0;
    const uint32_t ks = wbase + __lane_id();  // = k
    if (ks < n) {
        // caller order, or the sorted slot together with the particle itself
        const float4 ps = pos[ks];
        const float ox = (float)(-c.e24 * (double)acc.ax.value());
        const float oy = (float)(-c.e24 * (double)acc.ay.value());
        const float oz = (float)(-c.e24 * (double)acc.az.value());
        const double ei = c.e4 * (double)acc.ae.value();
        if constexpr (OUT == 2) {
            store_staged(&stage[__float_as_uint(ps.w)], ox, oy, oz, (float)ei);
        } else {
            constexpr bool NT = OUT == 1;
            const bool sorted = so.x != nullptr;  // uniform
            const uint32_t i = sorted ? ks : __float_as_uint(ps.w);
            if (sorted) {
                store_out<NT>(&so.x[ks], ps.x);
                store_out<NT>(&so.y[ks], ps.y);
                store_out<NT>(&so.z[ks], ps.z);
                if (so.perm) so.perm[ks] = __float_as_uint(ps.w);
            }
            store_out<NT>(&fx[i], ox);
            store_out<NT>(&fy[i], oy);
            store_out<NT>(&fz[i], oz);
            if (e) store_out<NT>(&e[i], (float)ei);
        }
        esum = ei;
        vsum = c.e24 * (double)acc.av.value();
    }
    block_sum2(esum, vsum);
    const unsigned long long bp = block_sum_u64(np);
    if (threadIdx.x == 0) {
        partials[2 * blk] = esum;
        partials[2 * blk + 1] = vsum;
        if (npairs && bp) atomicAdd(npairs, bp);
    }
}

__global__ __launch_bounds__(256) void k_split_outputs(const float4* __restrict__ stage, uint32_t n,
                                                       float* __restrict__ fx, float* __restrict__ fy,
                                                       float* __restrict__ fz, float* __restrict__ e) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(stage) + i);
    fx[i] = v.x;
    fy[i] = v.y;
    fz[i] = v.z;
    if (e) e[i] = v.w;
}

__global__ __launch_bounds__(256) void k_coarse_counts(const uint32_t* __restrict__ foff, GridGeom fg, GridGeom cg, int Kx,
                                                       int K, uint32_t* __restrict__ ccount) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cg.ncells) return;
    ccount[c] = coarse_count(foff, fg, cg, Kx, K, c);
}

// FP64 reduction of the per-block partials in a fixed order: stage 1 sums chunks of
// 1024 pairs into out (one block per chunk), stage 2 (one block, res != NULL) sums
// the chunk results into res.
__global__ __launch_bounds__(256) void k_finalize(const double* __restrict__ partials, uint32_t nblocks,
                                                  double* __restrict__ out, lj_result* __restrict__ res) {
    double a = 0.0, b = 0.0;
    const uint32_t lo = res ? 0 : blockIdx.x * 1024u;
    const uint32_t hi = res ? nblocks : (lo + 1024u < nblocks ? lo + 1024u : nblocks);
    for (uint32_t q = lo + threadIdx.x; q < hi; q += 256) {
        a += partials[2 * q];
        b += partials[2 * q + 1];
    }
    block_sum2(a, b);
    if (threadIdx.x == 0) {
        if (res) {
            res->e_total = 0.5 * a;
            res->virial = 0.5 * b;
        } else {
            out[2 * blockIdx.x] = a;
            out[2 * blockIdx.x + 1] = b;
        }
    }
}

// ------------------------------------------------------------------ pair lists
// Pass 1: number of accepted partners of each particle, stored by original index (a
// partner with a multiplicity above one is listed once).
template <int PRED>
__global__ __launch_bounds__(256) void k_pair_count(const float4* __restrict__ pos, const uint32_t* __restrict__ cell_of,
                                                    const uint32_t* __restrict__ off, uint32_t n, GridGeom g,
                                                    PairConst c, uint32_t* __restrict__ rowcnt) {
    const uint32_t k = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
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
                if (t != k && pair_mult<PRED>(c, pi, pj, d2) > 0) ++cnt;
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
    const uint32_t k = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
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
                if (t != k && pair_mult<PRED>(c, pi, pj, d2) > 0) cols[w++] = __float_as_uint(pj.w);
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
void launch_force(int prec, bool vir, const Shape& sh, unsigned nb, hipStream_t s, const float4* pos,
                  const uint32_t* cell_of, const uint32_t* off, uint32_t n, const GridGeom& g, const Trim& tr,
                  const PairConst& c, float* fx, float* fy, float* fz, float* e, double* partials,
                  unsigned long long* npairs, const SortedOut& so, float4* stage) {
    const size_t lds = sh.lds_bytes();
    const int out = stage ? 2 : non_temporal_outputs(n) ? 1 : 0;
#define LJ_LAUNCH_QN(PREC, VIR, Q, OUT)                                                                              \
    hipLaunchKernelGGL((k_force<PRED, PREC, VIR, Q, OUT>), dim3(nb), dim3(sh.block), lds, s, pos, cell_of, off, n, g, \
                       sh.refine_x, sh.refine, tr, c, (uint32_t)sh.depth, fx, fy, fz, e, partials, npairs, so, stage)
#define LJ_LAUNCH_Q(PREC, VIR, Q)                      \
    do {                                               \
        if (out == 2) LJ_LAUNCH_QN(PREC, VIR, Q, 2);   \
        else if (out == 1) LJ_LAUNCH_QN(PREC, VIR, Q, 1); \
        else LJ_LAUNCH_QN(PREC, VIR, Q, 0);            \
    } while (0)
#define LJ_LAUNCH(PREC, VIR)                   \
    do {                                       \
        if (sh.queued()) LJ_LAUNCH_Q(PREC, VIR, true); \
        else LJ_LAUNCH_Q(PREC, VIR, false);    \
    } while (0)
    switch (prec) {
    case LJ_PREC_MIXED: if (vir) LJ_LAUNCH(LJ_PREC_MIXED, true); else LJ_LAUNCH(LJ_PREC_MIXED, false); break;
    case LJ_PREC_FP32: if (vir) LJ_LAUNCH(LJ_PREC_FP32, true); else LJ_LAUNCH(LJ_PREC_FP32, false); break;
    default: if (vir) LJ_LAUNCH(LJ_PREC_FP32_COMP, true); else LJ_LAUNCH(LJ_PREC_FP32_COMP, false); break;
    }
#undef LJ_LAUNCH
#undef LJ_LAUNCH_Q
#undef LJ_LAUNCH_QN
}

}  // namespace

void split_outputs(const float4* stage, uint32_t n, float* fx, float* fy, float* fz, float* e, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_split_outputs, dim3(div_up(n, 256)), dim3(256), 0, s, stage, n, fx, fy, fz, e);
}

void finalize_energy(double* partials, unsigned nblocks, lj_result* res, hipStream_t s) {
    if (nblocks > 1024) {
        // stage 1 writes behind the nblocks input pairs (the partials buffer holds room)
        const unsigned chunks = (nblocks + 1023) / 1024;
        double* tmp = partials + 2 * (size_t)nblocks;
        hipLaunchKernelGGL(k_finalize, dim3(chunks), dim3(256), 0, s, (const double*)partials, (uint32_t)nblocks, tmp,
                           (lj_result*)nullptr);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, (const double*)tmp, (uint32_t)chunks, (double*)nullptr,
                           res);
    } else {
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, (const double*)partials, (uint32_t)nblocks,
                           (double*)nullptr, res);
    }
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

// Force grid of edge cw/Kx along x and cw/K along y and z (powers of two: exact),
// G' = (int)(L / edge) per axis.
int refine_geom(lj_context* ctx, const GridGeom& g, const lj_params* p, int Kx, int K, GridGeom* f) {
    *f = g;
    f->cw[0] = g.cw[0] / (float)Kx;
    f->cw[1] = g.cw[1] / (float)K;
    f->cw[2] = g.cw[2] / (float)K;
    uint64_t nc = 1;
    for (int d = 0; d < 3; ++d) {
        const float L = p->box_hi[d] - p->box_lo[d];
        const float q = L / f->cw[d];
        if (!(q < 16777216.0f)) return fail(ctx, LJ_ERR_OVERFLOW, "refined grid: more than 2^24 cells along one axis");
        int32_t gd = (int32_t)q;
        if (gd < 1) gd = 1;
        f->g[d] = gd;
        f->gmax[d] = (float)(gd - 1);
        nc *= (uint64_t)gd;
    }
    if (nc > 0x7fffffffull) return fail(ctx, LJ_ERR_OVERFLOW, "refined grid holds more than 2^31-1 cells");
    f->ncells = (uint32_t)nc;
    return LJ_OK;
}

// Shape of a call: explicit fields of p win.  Default grid: cells of rc/2 along y and
// z and rc/4 along x for large dense inputs (>= 8 particles per cell_edge cell and
// >= 2^22 particles, where fewer candidates pay for the larger key and cell arrays),
// the cell_edge grid otherwise (small inputs, and sparse ones whose extra empty cells
// cost more in the scan than they save; measured on MI355X at 0.1M-16.7M particles).
// The CUDALJ mode keeps the reference's 27 cells.  Block 256 and no queue.
Shape choose_shape(const lj_params* p, int64_t n, uint64_t ncells) {
    Shape sh;
    if (p->predicate == LJ_PRED_CUDALJ) {
        sh.refine = sh.refine_x = 1;
    } else {
        const double ppc = (double)n / (double)(ncells ? ncells : 1);
        const bool dense = ppc >= 8.0 && n >= ((int64_t)1 << 22) && (double)ncells * 16.0 < (double)(1u << 30);
        sh.refine = (p->refine == 1 || p->refine == 2 || p->refine == 4) ? p->refine : (dense ? 2 : 1);
        sh.refine_x = p->refine_x ? p->refine_x : (p->refine ? sh.refine : (dense ? 4 : 1));
        if (sh.refine_x < sh.refine) sh.refine_x = sh.refine;
    }
    sh.block = p->block ? p->block : 256;
    sh.depth = p->depth ? p->depth : 1;
    return sh;
}

Trim make_trim(const lj_params* p, const GridGeom& g) {
    Trim t;
    const double r = (double)p->rc * (1.0 + 0x1p-16);
    t.r2 = std::nextafter((float)(r * r), INFINITY);
    double mag = 0.0;
    for (int d = 0; d < 3; ++d) {
        mag = std::fmax(mag, std::fabs((double)p->box_lo[d]));
        mag = std::fmax(mag, std::fabs((double)p->box_hi[d]));
        t.hi[d] = p->box_hi[d];
    }
    // 2^-19 mag bounds the rounding of v - lo and (times the cell edge) the few ulp
    // by which (v - lo) inv_cw -+ w inv_cw and the cell coordinate (v - lo) / cw differ
    // from their exact values; the rc term covers the square root of the window
    const double cwmax = std::fmax((double)g.cw[0], std::fmax((double)g.cw[1], (double)g.cw[2]));
    t.margin = (float)(0x1p-19 * mag + (double)p->rc * 0x1p-14 + 8.0 * 0x1p-24 * cwmax);
    t.inv_cw = 1.0f / g.cw[0];
    t.on = p->predicate != LJ_PRED_CUDALJ;
    return t;
}

int compute_cellgrid(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
                     const GridGeom& g, const PairConst& c, float* fx, float* fy, float* fz, float* e,
                     lj_result* result, hipStream_t s, const SortedOut& so) {
    const Shape sh = choose_shape(p, n, g.ncells);
    const bool refined = sh.refine > 1 || sh.refine_x > 1;
    GridGeom fg = g;
    if (refined) LJ_TRY(refine_geom(ctx, g, p, sh.refine_x, sh.refine, &fg));
    const Trim tr = make_trim(p, fg);
    {
        PhaseTimer t(ctx, LJ_PHASE_BIN, s);
        LJ_TRY(bin_particles(ctx, x, y, z, n, fg, s));
    }
    lj_result* res;
    LJ_TRY(result_slot(ctx, result, &res));
    {
        PhaseTimer t(ctx, LJ_PHASE_OTHER, s);
        if (!refined) {
            LJ_TRY(count_candidates(ctx, g, &res->n_candidates, s));
        } else {
            // candidate count on the cell_edge grid (ComputeNbInteractions semantics)
            BinWork& w = ctx->bin;
            LJ_HIP(ctx, w.ccounts.grow((size_t)g.ncells * 4));
            LJ_HIP(ctx, w.coffsets.grow(((size_t)g.ncells + 1) * 4));
            hipLaunchKernelGGL(k_coarse_counts, dim3(div_up(g.ncells, 256)), dim3(256), 0, s,
                               (const uint32_t*)w.offsets.as<uint32_t>(), fg, g, sh.refine_x, sh.refine,
                               w.ccounts.as<uint32_t>());
            uint32_t* coff = w.coffsets.as<uint32_t>();
            LJ_TRY(scan_u32(ctx, w.ccounts.as<uint32_t>(), coff, g.ncells, coff + g.ncells, s));
            LJ_TRY(count_candidates_from(ctx, g, coff, &res->n_candidates, s));
        }
        fill_u64(&res->n_pairs, 1, 0ull, s);
    }
    const unsigned nb = div_up(n > 0 ? n : 1, sh.block);
    LJ_HIP(ctx, ctx->bin.partials.grow(((size_t)nb + nb / 1024 + 1) * 2 * sizeof(double)));
    double* partials = ctx->bin.partials.as<double>();
    const int prec = p->precision;
    const bool vir = !(p->flags & LJ_FLAG_NO_VIRIAL);
    const float4* pos = ctx->bin.pos.as<float4>();
    const uint32_t* cell_of = ctx->bin.sorted_keys();
    const uint32_t* off = ctx->bin.offsets.as<uint32_t>();
    const uint32_t nn = (uint32_t)n;
    // large outputs at shuffled indices go through the staging array
    float4* stage = nullptr;
    if (non_temporal_outputs(n) && !so.x) {
        LJ_HIP(ctx, ctx->bin.stage.grow((size_t)n * sizeof(float4)));
        stage = ctx->bin.stage.as<float4>();
    }
    {
        PhaseTimer tf(ctx, LJ_PHASE_FORCE, s);
#define LJ_PRED_CASE(P)                                                                                         \
    case P:                                                                                                     \
        launch_force<P>(prec, vir, sh, nb, s, pos, cell_of, off, nn, fg, tr, c, fx, fy, fz, e, partials,      \
                        &res->n_pairs, so, stage);                                                              \
        break;
        switch (p->predicate) {
            LJ_PRED_CASE(LJ_PRED_AABB)
            LJ_PRED_CASE(LJ_PRED_CUSTOM)
            LJ_PRED_CASE(LJ_PRED_CUDALJ)
            LJ_PRED_CASE(LJ_PRED_SPHERE)
            LJ_PRED_CASE(LJ_PRED_TRIANGLE)
        }
#undef LJ_PRED_CASE
    }
    PhaseTimer to(ctx, LJ_PHASE_OTHER, s);
    if (stage) split_outputs(stage, nn, fx, fy, fz, e, s);
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

int lj_compute_sorted(lj_context* ctx, float* x, float* y, float* z, int64_t n, const lj_params* p, float* fx,
                      float* fy, float* fz, float* e, uint32_t* perm, lj_result* result, void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    GridGeom g;
    PairConst c;
    LJ_TRY(make_geom(ctx, p, n, &g, &c));
    if (n > 0 && (!x || !y || !z || !fx || !fy || !fz))
        return fail(ctx, LJ_ERR_INVALID_ARG, "positions and forces must not be NULL");
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    // the sort reads x, y, z before the force launch overwrites them
    SortedOut so;
    so.x = x;
    so.y = y;
    so.z = z;
    so.perm = perm;
    return compute_cellgrid(ctx, x, y, z, n, p, g, c, fx, fy, fz, e, result, (hipStream_t)stream, so);
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
    if (perm) LJ_TRY(extract_perm(ctx, n, perm, s));
    return LJ_OK;
}

}  // extern "C"
