// lj_stencil.h -- the cell-stencil pair loop shared by the single-system and the
// batched force kernels, and the host-side choices that shape it.
//
// Reference: ComputeParticleInterationsParPartNoLoop (Cuda-version/cudaLJ.cu:299-341)
// visits, for one particle, the particles of the clipped 27 neighbour cells in z -> y
// -> x order and skips itself.  Cell order is z-major (cudaLJ.cu:127-134), so the
// x-cells of one (y, z) row of the stencil are one contiguous slot range.  The rows are
// visited here in another order (mirror pairs, see stencil_pairs): the pair set is the
// same, and the FP64 sums differ only in rounding.
#pragma once

#include "lj_internal.h"
#include "lj_pair.h"

namespace lj {

// Stencil trimming constants: a row of the stencil is skipped when its (y,z) distance
// to particle i exceeds the cutoff, and the x-cells of a row are limited to the cells
// of [x_i - w, x_i + w] with w = sqrt(rc^2 - gap^2), found in cell units as
// (x_i - lo) inv_cw -+ w inv_cw with the rounded reciprocal inv_cw of the cell edge.
// Every bound is widened by `margin`, which exceeds the FP32 rounding of these steps
// and their difference from the cell assignment's correctly rounded quotient (a few
// ulp of the cell coordinate), so a row or cell is dropped only when no particle in
// it can pass the predicate.
struct Trim {
    float r2;       // (rc (1 + 2^-16))^2, rounded up
    float margin;   // absolute widening of every bound
    float hi[3];    // box_hi: pruning applies to particles inside the box
    float inv_cw;   // 1 / x cell edge of the force grid, rounded
    int on;         // 0 for the CUDALJ mode (no cutoff)
};

// Kernel shape of a force launch: force-grid refinement (cells of edge cell_edge/Kx
// along x and cell_edge/K along y and z; the stencil reaches Kx cells along x and K
// along y and z), particles per workgroup and LDS queue entries per particle (depth 1:
// no queue, the pair arithmetic runs in the candidate loop).
struct Shape {
    int refine, refine_x, block, depth;
    bool queued() const { return depth > 1; }
    size_t lds_bytes() const { return queued() ? (size_t)depth * (size_t)block * 16 : 0; }
};

constexpr int kMaxBlock = 256;
// candidates loaded and tested per iteration of the stencil's row loop (the sorted
// positions carry kCandidates - 1 spare entries, so the loads need no clamp)
constexpr int kCandidates = 3;
static_assert(kCandidates - 1 <= kPosSpare, "unclamped candidate loads need spare slots");
constexpr int kMaxDepth = 32;

__device__ inline void cell_coords(const GridGeom& g, uint32_t key, int32_t& cx, int32_t& cy, int32_t& cz) {
    cx = (int32_t)(key % (uint32_t)g.g[0]);
    const uint32_t r = key / (uint32_t)g.g[0];
    cy = (int32_t)(r % (uint32_t)g.g[1]);
    cz = (int32_t)(r / (uint32_t)g.g[1]);
}

// Particles in cell c of the cell_edge grid cg, from the offsets foff of the refined
// grid fg (edge cg.cw / Kx along x, cg.cw / K along y and z).  For a power of two K,
// x / (cw/K) is exactly K (x / cw) in binary floating point and floor(floor(K q) / K) =
// floor(q), so coarse cell X holds fine cells KX .. KX+K-1 (and the extra last fine
// cells, which the coarse clamp also maps to the last X).
__device__ inline uint32_t coarse_count(const uint32_t* __restrict__ foff, const GridGeom& fg, const GridGeom& cg, int Kx,
                                        int K, uint32_t c) {
    const int32_t X = (int32_t)(c % (uint32_t)cg.g[0]);
    const uint32_t rem = c / (uint32_t)cg.g[0];
    const int32_t Y = (int32_t)(rem % (uint32_t)cg.g[1]), Z = (int32_t)(rem / (uint32_t)cg.g[1]);
    const int32_t fx0 = Kx * X, fx1 = X == cg.g[0] - 1 ? fg.g[0] - 1 : Kx * X + Kx - 1;
    const int32_t fy0 = K * Y, fy1 = Y == cg.g[1] - 1 ? fg.g[1] - 1 : K * Y + K - 1;
    const int32_t fz0 = K * Z, fz1 = Z == cg.g[2] - 1 ? fg.g[2] - 1 : K * Z + K - 1;
    uint32_t s = 0;
    for (int32_t zz = fz0; zz <= fz1; ++zz)
        for (int32_t yy = fy0; yy <= fy1; ++yy) {
            const uint32_t row = (uint32_t)((zz * fg.g[1] + yy) * fg.g[0]);
            s += foff[row + fx1 + 1] - foff[row + fx0];
        }
    return s;
}

// Accepted partners of the particle in sorted slot k (position pi, cell `key` of grid
// g, whose cell 0 is global cell `cbase`), summed into acc; returns their number.
// The stencil is (2Rx+1) x (2R+1)^2 cells of g, visited as (2R+1)^2 x-rows in mirror
// pairs (below), rows and x-cells beyond the cutoff trimmed.
//
// Pairs.  With QUEUE = false (the default shape) the FP64 pair arithmetic of an
// accepted candidate runs inside the candidate loop; it then runs for the whole wave
// whenever any lane accepts, with about a quarter of the lanes busy.  With QUEUE =
// true the candidate loop only evaluates the predicate and appends an accepted
// partner's position to the lane's ring of `depth` entries in LDS.  The rings of a
// wave are one block of depth x 64 x 16 bytes (entry q of lane l at (q 64 + l) 16), so
// an entry is one 128-bit access at a shifted address and the lanes of a wave use
// consecutive words.  A
// drain step pops the oldest entry of every lane that has one and runs the pair
// arithmetic on it; each candidate iteration drains once when at least kDrainLanes
// lanes hold entries, and up to kCandidates times when a ring could not take the
// iteration's candidates.  Entries are popped in the order they were pushed, so the
// sums and their bits are those of QUEUE = false.  Every lane of the wave must call
// this (inactive lanes with active = false), because the drains are wave-wide.
constexpr int kDrainLanes = 40;

struct __align__(16) QEntry {
    float x, y, z, pad;
};

template <int PRED, int PREC, bool VIR, bool QUEUE>
__device__ inline uint32_t stencil_pairs(const float4* __restrict__ pos, const uint32_t* __restrict__ off, uint32_t k,
                                         bool active, const float4& pi, const GridGeom& g, uint32_t cbase,
                                         uint32_t key, int Rx, int R, const Trim& tr, const PairConst& c,
                                         QEntry* __restrict__ queue, unsigned depth,
                                         PairSums<PREC, PRED == LJ_PRED_CUDALJ, VIR>& acc) {
    // only the TRIANGLE ray model counts a pair more than once (queued partners then
    // get their multiplicity again when drained)
    constexpr bool kMultiple = PRED == LJ_PRED_TRIANGLE;
    unsigned pc = 0, head = 0;  // entries queued by this lane, index of the oldest
    uint32_t np = 0;
    auto push = [&](const float4& pj) {
        queue[((head + pc) & (depth - 1)) * kWaveSize] = QEntry{pj.x, pj.y, pj.z, 0.f};
        ++pc;
    };
    auto drain_one = [&]() {
        if (pc > 0) {
            const QEntry q = queue[head * kWaveSize];
            const float4 pj = make_float4(q.x, q.y, q.z, 0.f);
            const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const float d2 = dist2(dx, dy, dz);
            acc.add(c, pj, dx, dy, dz, d2, kMultiple ? pair_mult<PRED>(c, pi, pj, d2) : 1);
            head = (head + 1) & (depth - 1);
            --pc;
        }
    };
    auto visit = [&](uint32_t t, const float4& pj) {
        const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const float d2 = dist2(dx, dy, dz);
        // the cutoff predicates reject d2 = 0, hence the particle itself; the CUDALJ
        // mode accepts every candidate and skips itself by slot
        if constexpr (PRED == LJ_PRED_CUDALJ) {
            if (t == k) return;
        }
        const int mult = pair_mult<PRED>(c, pi, pj, d2);
        if (!mult) return;
        np += (uint32_t)mult;
        if constexpr (QUEUE) {
            push(pj);
        } else {
            acc.add(c, pj, dx, dy, dz, d2, mult);
        }
    };
    int32_t cx = 0, cy = 0, cz = 0;
    if (active) cell_coords(g, key, cx, cy, cz);
    const bool trim = active && tr.on && pi.x >= g.lo[0] && pi.y >= g.lo[1] && pi.z >= g.lo[2] &&
                      pi.x <= tr.hi[0] && pi.y <= tr.hi[1] && pi.z <= tr.hi[2];
    const int32_t x0 = cx - Rx > 0 ? cx - Rx : 0, x1 = cx + Rx < g.g[0] - 1 ? cx + Rx : g.g[0] - 1;
    // an inactive lane has an empty stencil (z0 > z1) but still takes part in the drains
    const int32_t z0 = active ? (cz - R > 0 ? cz - R : 0) : 1;
    const int32_t z1 = active ? (cz + R < g.g[2] - 1 ? cz + R : g.g[2] - 1) : 0;
    const int32_t y0 = cy - R > 0 ? cy - R : 0, y1 = cy + R < g.g[1] - 1 ? cy + R : g.g[1] - 1;
    // x cell coordinate of particle i by the reciprocal (trimming only): the x-cells of
    // a row are those of qx -+ w inv_cw, truncated (a negative lower bound truncates to
    // 0, like the clamp of the cell rule) and clipped to [x0, x1]
    const float qx = (pi.x - g.lo[0]) * tr.inv_cw;
    // Per particle, once: its distances to the faces of its own cell in y and z, and
    // its own row.  The gap to the row d cells away is then |d| - 1 cell edges plus the
    // distance to the near face (one FMA), and a row's index is the own row plus a
    // stencil-uniform offset; both were per-row arithmetic on the cell coordinates.
    // (The FMA rounds differently from the direct lo + c cw - v, by a few ulp of the
    // coordinates, far inside the margin.)
    float fy_lo = 0.f, fy_hi = 0.f, fz_lo = 0.f, fz_hi = 0.f;
    if (trim) {
        fy_lo = pi.y - (g.lo[1] + (float)cy * g.cw[1]);
        fy_hi = (g.lo[1] + (float)(cy + 1) * g.cw[1]) - pi.y;
        fz_lo = pi.z - (g.lo[2] + (float)cz * g.cw[2]);
        fz_hi = (g.lo[2] + (float)(cz + 1) * g.cw[2]) - pi.z;
    }
    const uint32_t home = cbase + (uint32_t)((cz * g.g[1] + cy) * g.g[0]);
    // the clipped stencil as offsets from the particle's cell
    const int32_t dz0 = z0 - cz, dz1 = z1 - cz, dy0 = y0 - cy, dy1 = y1 - cy;
    auto gap = [&](int32_t d, float near_lo, float near_hi, float cw) {
        float v = 0.f;
        if (d > 0) v = __builtin_fmaf((float)(d - 1), cw, near_hi);
        else if (d < 0) v = __builtin_fmaf((float)(-d - 1), cw, near_lo);
        return fmaxf(v - tr.margin, 0.f);
    };
    // Slot range [b, en) of the trimmed x-cells of row (cz + dz, cy + dy); empty
    // outside the clipped stencil or beyond the cutoff.
    auto row_range = [&](int32_t dz, int32_t dy, uint32_t& b, uint32_t& en) {
        b = en = 0u;
        if (dz < dz0 || dz > dz1 || dy < dy0 || dy > dy1) return;
        int32_t xa = x0, xb = x1;
        if (trim) {
            const float gz = gap(dz, fz_lo, fz_hi, g.cw[2]);
            const float gy = gap(dy, fy_lo, fy_hi, g.cw[1]);
            const float w2 = tr.r2 - gz * gz - gy * gy;
            if (w2 < 0.f) return;
            // hardware square root (about 1 ulp, no correctly rounded fix-up): its
            // error is far inside the margin, which exceeds rc 2^-14
            const float w = __builtin_amdgcn_sqrtf(w2) + tr.margin;
            xa = min(max(x0, (int32_t)__builtin_fmaf(-w, tr.inv_cw, qx)), x1);
            xb = min(x1, (int32_t)__builtin_fmaf(w, tr.inv_cw, qx));
        }
        const uint32_t row = home + (uint32_t)(dz * g.g[1] + dy) * (uint32_t)g.g[0];
        b = off[row + xa];
        en = off[row + xb + 1];
    };
    // Rows go in mirror pairs, (cz + dz, cy + dy) and then (cz - dz, cy - dy) for
    // (dz, dy) in one half of the stencil, each pair walked as one stream of
    // candidates.  A row and its mirror hold about as many candidates together wherever
    // the particle sits in its cell, so the lanes of a wave finish a pair at closer
    // times than a single row, and a wave runs each step for its longest lane.
    for (int32_t dz = 0; dz <= R; ++dz)
        for (int32_t dy = dz == 0 ? 0 : -R; dy <= R; ++dy) {
            if constexpr (QUEUE) {  // every lane takes part here
                while (__popcll(__ballot(pc > 0)) >= (unsigned)kDrainLanes) drain_one();
            }
            uint32_t t, en, b2, e2;
            row_range(dz, dy, t, en);
            if (dz == 0 && dy == 0) b2 = e2 = 0u;
            else row_range(-dz, -dy, b2, e2);
            if (t >= en) {
                t = b2;
                en = e2;
                b2 = e2;
            }
            // candidate t + u exists iff t + u < en, i.e. t < enu[u] = max(en - u, 0)
            uint32_t enu[kCandidates];
#pragma unroll
            for (int u = 1; u < kCandidates; ++u) enu[u] = en > (uint32_t)u ? en - (uint32_t)u : 0u;
            while (t < en) {
                if constexpr (QUEUE) {
                    // room for this iteration's candidates in every walking lane's ring
                    // (a lane pushes at most kCandidates per iteration), then one step
                    // when enough lanes are waiting
#pragma unroll
                    for (int u = 0; u < kCandidates; ++u)
                        if (__any(pc + kCandidates > depth)) drain_one();
                    if (__popcll(__ballot(pc > 0)) >= (unsigned)kDrainLanes) drain_one();
                }
                // slots up to t + kCandidates - 1 <= n + kCandidates - 2 exist (pos has
                // spare entries); they count only below en.  One 64-bit base, the
                // candidates at immediate offsets from it.
                const float4* pt = pos + t;
                float4 pj[kCandidates];
#pragma unroll
                for (int u = 0; u < kCandidates; ++u) pj[u] = pt[u];
                visit(t, pj[0]);
#pragma unroll
                for (int u = 1; u < kCandidates; ++u)
                    if (t < enu[u]) visit(t + u, pj[u]);
                t += kCandidates;
                if (t >= en) {  // on to the mirror row (then t = en = e2 ends the pair)
                    t = b2;
                    en = e2;
                    b2 = e2;
#pragma unroll
                    for (int u = 1; u < kCandidates; ++u) enu[u] = en > (uint32_t)u ? en - (uint32_t)u : 0u;
                }
            }
        }
    // rows end at different times in different lanes: drain while any ring holds pairs
    if constexpr (QUEUE)
        while (__any(pc > 0)) drain_one();
    return np;
}

// Host side (force.hip).
int refine_geom(lj_context* ctx, const GridGeom& g, const lj_params* p, int Kx, int K, GridGeom* f);
Trim make_trim(const lj_params* p, const GridGeom& g);
// Shape of a call over n particles in ncells cells of the cell_edge grid: explicit
// fields of p (validated by make_geom) win; zero fields take the defaults.
Shape choose_shape(const lj_params* p, int64_t n, uint64_t ncells);

}  // namespace lj
