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
#include <cmath>

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

// Stencil trimming constants: a row of the stencil is skipped when its (y,z) distance
// to particle i exceeds the cutoff, and the x-cells of a row are limited to
// [x_i - w, x_i + w] with w = sqrt(rc^2 - gap^2).  Every bound is widened by `margin`,
// which exceeds the FP32 rounding of a cell assignment and of x_i +- w, so a row or
// cell is dropped only when no particle in it can pass the predicate.
struct Trim {
    float r2;       // (rc (1 + 2^-16))^2, rounded up
    float margin;   // absolute widening of every bound
    float hi[3];    // box_hi: pruning applies to particles inside the box
    int on;         // 0 for the CUDALJ mode (no cutoff)
};

// Lower bound of the distance along one axis from coordinate v to cell cc.
__device__ inline float axis_gap(float v, int32_t cc, int32_t home, float lo, float cw, float margin) {
    float gap = 0.f;
    if (cc > home) gap = (lo + (float)cc * cw) - v;
    else if (cc < home) gap = v - (lo + (float)(cc + 1) * cw);
    return fmaxf(gap - margin, 0.f);
}

// One thread per sorted slot k.  The stencil is (2R+1)^3 cells of the force grid
// (R = 1: the 27 cells of cudaLJ.cu:319-321 on the cell_edge grid; R = 2 or 4: the
// half- or quarter-edge grid), visited as (2R+1)^2 x-rows in z -> y order, each row
// a contiguous slot range; rows and x-cells beyond the cutoff are trimmed (Trim).
// Each row is read two candidates per step, so two position loads are in flight.
template <int PRED, int PREC, bool VIR>
__global__ __launch_bounds__(256) void k_force_simple(const float4* __restrict__ pos, const uint32_t* __restrict__ cell_of,
                                                      const uint32_t* __restrict__ off, uint32_t n, GridGeom g, int R,
                                                      Trim tr, PairConst c, float* __restrict__ fx, float* __restrict__ fy,
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
        const bool trim = tr.on && pi.x >= g.lo[0] && pi.y >= g.lo[1] && pi.z >= g.lo[2] && pi.x <= tr.hi[0] &&
                          pi.y <= tr.hi[1] && pi.z <= tr.hi[2];
        const int32_t x0 = cx - R > 0 ? cx - R : 0, x1 = cx + R < g.g[0] - 1 ? cx + R : g.g[0] - 1;
        const int32_t z0 = cz - R > 0 ? cz - R : 0, z1 = cz + R < g.g[2] - 1 ? cz + R : g.g[2] - 1;
        const int32_t y0 = cy - R > 0 ? cy - R : 0, y1 = cy + R < g.g[1] - 1 ? cy + R : g.g[1] - 1;
        auto visit = [&](uint32_t t, float4 pj) {
            const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const float d2 = dist2(dx, dy, dz);
            if (t == k || !accept<PRED>(c, d2)) return;
            ++np;
            acc.add(c, pj, dx, dy, dz, d2);
        };
        for (int32_t zz = z0; zz <= z1; ++zz) {
            const float gz = trim ? axis_gap(pi.z, zz, cz, g.lo[2], g.cw, tr.margin) : 0.f;
            for (int32_t yy = y0; yy <= y1; ++yy) {
                int32_t xa = x0, xb = x1;
                if (trim) {
                    const float gy = axis_gap(pi.y, yy, cy, g.lo[1], g.cw, tr.margin);
                    const float w2 = tr.r2 - gz * gz - gy * gy;
                    if (w2 < 0.f) continue;
                    const float w = __builtin_sqrtf(w2) + tr.margin;
                    xa = max(x0, axis_cell(pi.x - w, g.lo[0], g.cw, g.gmax[0]));
                    xb = min(x1, axis_cell(pi.x + w, g.lo[0], g.cw, g.gmax[0]));
                }
                const uint32_t row = (uint32_t)((zz * g.g[1] + yy) * g.g[0]);
                const uint32_t b = off[row + xa], en = off[row + xb + 1];
                uint32_t t = b;
                for (; t + 1 < en; t += 2) {
                    const float4 p0 = pos[t], p1 = pos[t + 1];
                    visit(t, p0);
                    visit(t + 1, p1);
                }
                if (t < en) visit(t, pos[t]);
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

// Counts of the cell_edge grid from the refined grid's offsets.  For a refinement by
// K = 2 or 4, x / (cw/K) is exactly K (x / cw) in binary floating point and
// floor(floor(K q) / K) = floor(q), so coarse cell X holds fine cells KX .. KX+K-1 (and
// the extra last fine cells, which the coarse clamp also maps to the last X).
__global__ __launch_bounds__(256) void k_coarse_counts(const uint32_t* __restrict__ foff, GridGeom fg, GridGeom cg, int K,
                                                       uint32_t* __restrict__ ccount) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cg.ncells) return;
    const int32_t X = (int32_t)(c % (uint32_t)cg.g[0]);
    const uint32_t rem = c / (uint32_t)cg.g[0];
    const int32_t Y = (int32_t)(rem % (uint32_t)cg.g[1]), Z = (int32_t)(rem / (uint32_t)cg.g[1]);
    const int32_t fx0 = K * X, fx1 = X == cg.g[0] - 1 ? fg.g[0] - 1 : K * X + K - 1;
    const int32_t fy0 = K * Y, fy1 = Y == cg.g[1] - 1 ? fg.g[1] - 1 : K * Y + K - 1;
    const int32_t fz0 = K * Z, fz1 = Z == cg.g[2] - 1 ? fg.g[2] - 1 : K * Z + K - 1;
    uint32_t s = 0;
    for (int32_t zz = fz0; zz <= fz1; ++zz)
        for (int32_t yy = fy0; yy <= fy1; ++yy) {
            const uint32_t row = (uint32_t)((zz * fg.g[1] + yy) * fg.g[0]);
            s += foff[row + fx1 + 1] - foff[row + fx0];
        }
    ccount[c] = s;
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
                  const uint32_t* off, uint32_t n, const GridGeom& g, int R, const Trim& tr, const PairConst& c, float* fx,
                  float* fy, float* fz, float* e, double* partials, unsigned long long* npairs) {
#define LJ_LAUNCH(PREC, VIR)                                                                                         \
    hipLaunchKernelGGL((k_force_simple<PRED, PREC, VIR>), grid, dim3(256), 0, s, pos, cell_of, off, n, g, R, tr, c, fx, \
                       fy, fz, e, partials, npairs)
    switch (prec) {
    case LJ_PREC_MIXED: if (vir) LJ_LAUNCH(LJ_PREC_MIXED, true); else LJ_LAUNCH(LJ_PREC_MIXED, false); break;
    case LJ_PREC_FP32: if (vir) LJ_LAUNCH(LJ_PREC_FP32, true); else LJ_LAUNCH(LJ_PREC_FP32, false); break;
    default: if (vir) LJ_LAUNCH(LJ_PREC_FP32_COMP, true); else LJ_LAUNCH(LJ_PREC_FP32_COMP, false); break;
    }
#undef LJ_LAUNCH
}

}  // namespace

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

// Force grid of edge cw/K (K = 2 or 4, exact), G' = (int)(L / (cw/K)).
static int refine_geom(lj_context* ctx, const GridGeom& g, const lj_params* p, int K, GridGeom* f) {
    *f = g;
    f->cw = g.cw / (float)K;
    uint64_t nc = 1;
    for (int d = 0; d < 3; ++d) {
        const float L = p->box_hi[d] - p->box_lo[d];
        int32_t gd = (int32_t)(L / f->cw);
        if (gd < 1) gd = 1;
        f->g[d] = gd;
        f->gmax[d] = (float)(gd - 1);
        nc *= (uint64_t)gd;
    }
    if (nc > 0x7fffffffull) return fail(ctx, LJ_ERR_OVERFLOW, "refined grid holds more than 2^31-1 cells");
    f->ncells = (uint32_t)nc;
    return LJ_OK;
}

// Force-grid refinement for a call: an explicit params->refine (1, 2 or 4) wins;
// otherwise the half-edge grid when the cell_edge grid averages >= 2 particles per
// cell and the refined grid stays within 2^28 cells.
static int choose_refine(const lj_params* p, int64_t n, const GridGeom& g) {
    if (p->predicate == LJ_PRED_CUDALJ) return 1;  // the reference stencil is the 27 cells
    if (p->refine == 1 || p->refine == 2 || p->refine == 4) return p->refine;
    const double ppc = (double)n / (double)g.ncells;
    return (ppc >= 2.0 && (double)g.ncells * 8.0 < (double)(1u << 28)) ? 2 : 1;
}

static Trim make_trim(const lj_params* p, const GridGeom& g) {
    Trim t;
    const double r = (double)p->rc * (1.0 + 0x1p-16);
    t.r2 = std::nextafter((float)(r * r), INFINITY);
    double mag = 0.0;
    for (int d = 0; d < 3; ++d) {
        mag = std::fmax(mag, std::fabs((double)p->box_lo[d]));
        mag = std::fmax(mag, std::fabs((double)p->box_hi[d]));
        t.hi[d] = p->box_hi[d];
    }
    t.margin = (float)(8.0 * 0x1p-24 * mag + (double)p->rc * 0x1p-14 + 8.0 * 0x1p-24 * (double)g.cw);
    t.on = p->predicate != LJ_PRED_CUDALJ;
    return t;
}

int compute_cellgrid(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n, const lj_params* p,
                     const GridGeom& g, const PairConst& c, float* fx, float* fy, float* fz, float* e,
                     lj_result* result, hipStream_t s) {
    const int refine = choose_refine(p, n, g);
    GridGeom fg = g;
    if (refine > 1) LJ_TRY(refine_geom(ctx, g, p, refine, &fg));
    const Trim tr = make_trim(p, fg);
    {
        PhaseTimer t(ctx, LJ_PHASE_BIN, s);
        LJ_TRY(bin_particles(ctx, x, y, z, n, fg, s));
    }
    lj_result* res;
    LJ_TRY(result_slot(ctx, result, &res));
    {
        PhaseTimer t(ctx, LJ_PHASE_OTHER, s);
        if (refine == 1) {
            LJ_TRY(count_candidates(ctx, g, &res->n_candidates, s));
        } else {
            // candidate count on the cell_edge grid (ComputeNbInteractions semantics)
            BinWork& w = ctx->bin;
            LJ_HIP(ctx, w.ccounts.grow((size_t)g.ncells * 4));
            LJ_HIP(ctx, w.coffsets.grow(((size_t)g.ncells + 1) * 4));
            hipLaunchKernelGGL(k_coarse_counts, dim3(div_up(g.ncells, 256)), dim3(256), 0, s,
                               (const uint32_t*)w.offsets.as<uint32_t>(), fg, g, refine, w.ccounts.as<uint32_t>());
            uint32_t* coff = w.coffsets.as<uint32_t>();
            LJ_TRY(scan_u32(ctx, w.ccounts.as<uint32_t>(), coff, g.ncells, coff + g.ncells, s));
            LJ_TRY(count_candidates_from(ctx, g, coff, &res->n_candidates, s));
        }
        fill_u64(&res->n_pairs, 1, 0ull, s);
    }
    const unsigned nb = div_up(n > 0 ? n : 1, 256);
    LJ_HIP(ctx, ctx->bin.partials.grow(((size_t)nb + nb / 1024 + 1) * 2 * sizeof(double)));
    double* partials = ctx->bin.partials.as<double>();
    const int prec = p->precision;
    const bool vir = !(p->flags & LJ_FLAG_NO_VIRIAL);
    const float4* pos = ctx->bin.pos.as<float4>();
    const uint32_t* cell_of = ctx->bin.sorted_keys();
    const uint32_t* off = ctx->bin.offsets.as<uint32_t>();
    const uint32_t nn = (uint32_t)n;
    {
        PhaseTimer tf(ctx, LJ_PHASE_FORCE, s);
#define LJ_PRED_CASE(P)                                                                                         \
    case P:                                                                                                     \
        launch_force<P>(prec, vir, dim3(nb), s, pos, cell_of, off, nn, fg, refine, tr, c, fx, fy, fz, e, partials, \
                        &res->n_pairs);                                                                         \
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
    if (perm) LJ_TRY(extract_perm(ctx, n, perm, s));
    return LJ_OK;
}

}  // extern "C"
