// lj_device.h -- arithmetic shared by every kernel of the engine.
//
// The cell key and the pair predicate are defined once here.  Binning, the
// cell-grid kernels, the pair-list kernels and the LBVH traversal all call these
// functions, so that every path accepts exactly the same pairs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lj_amd.h"

namespace lj {

// Kernel-side view of lj_params with the derived constants precomputed on the host.
struct GridGeom {
    float lo[3];
    float cw[3];       // cell edge per axis (equal on the cell_edge grid)
    int32_t g[3];      // cells per axis
    float gmax[3];     // (float)(g-1), the clamp of the cell coordinate
    uint32_t ncells;
};

struct PairConst {
    double sig2d;      // sigma*sigma, exact in FP64
    double softd;      // softening2 in CUDALJ mode, 0 otherwise (FP64)
    float rc2;         // rc*rc in FP32
    float sig2;        // sigma*sigma rounded to FP32
    float soft;        // softening2 in CUDALJ mode, 0 otherwise
    float tri_lo;      // largest float <= 1e-4 (TRIANGLE predicate)
    float rc;          // rc (TRIANGLE predicate compares sqrt(d2) < rc)
    double e4;         // 4*epsilon
    double e24;        // 24*epsilon
    int32_t pred;
};

// PosToCellCoord (Cuda-version/cudaLJ.cu:108-118): (int)(x / cw) in FP32 with IEEE
// division, here on x - lo and clamped to [0, G-1] (the largest float below the box
// edge rounds to G at rc = 0.04, 0.01 and 0.005).  fmaxf sends NaN to cell 0.
__host__ __device__ inline int32_t axis_cell(float x, float lo, float cw, float gmax) {
    float q = (x - lo) / cw;
    q = fmaxf(q, 0.0f);
    q = fminf(q, gmax);
    return (int32_t)q;
}

// CellCoordToIndex (cudaLJ.cu:127-134): ((z*Gy)+y)*Gx+x.
__host__ __device__ inline uint32_t cell_key(const GridGeom& g, float x, float y, float z) {
    const int32_t cx = axis_cell(x, g.lo[0], g.cw[0], g.gmax[0]);
    const int32_t cy = axis_cell(y, g.lo[1], g.cw[1], g.gmax[1]);
    const int32_t cz = axis_cell(z, g.lo[2], g.cw[2], g.gmax[2]);
    return (uint32_t)((cz * g.g[1] + cy) * g.g[0] + cx);
}

// Canonical squared distance on FP32 inputs, d = p_j - p_i.  Each difference of two
// FP32 values is one correctly rounded operation (relative error <= 2^-24), so the
// global coordinates lose nothing against cell-local ones.
__host__ __device__ inline float dist2(float dx, float dy, float dz) {
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Pair predicates; see lj_amd.h for the reference line of each.
template <int PRED>
__host__ __device__ inline bool accept(const PairConst& c, float d2) {
    if constexpr (PRED == LJ_PRED_AABB) return d2 <= c.rc2 && d2 >= 1e-16f;
    if constexpr (PRED == LJ_PRED_CUSTOM) return d2 < c.rc2 && d2 >= 1e-16f;
    if constexpr (PRED == LJ_PRED_CUDALJ) return true;  // j != i is decided by index
    if constexpr (PRED == LJ_PRED_SPHERE) return d2 < c.rc2 && d2 > 0.0f;
    if constexpr (PRED == LJ_PRED_TRIANGLE) {
        const float d = __builtin_sqrtf(d2);
        return d < c.rc && d > c.tri_lo;
    }
    return false;
}

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin (workgroup b runs on
// XCD b % 8), and each XCD has its own L2.  This bijection gives XCD x a contiguous
// run of logical blocks, so neighbouring blocks, which read the same neighbour cells,
// share an L2: the first nb % 8 XCDs take q + 1 blocks, the others q = nb / 8.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t xcd = b % 8u, idx = b / 8u, q = nb / 8u, r = nb % 8u;
    return xcd < r ? xcd * (q + 1u) + idx : r * (q + 1u) + (xcd - r) * q + idx;
}

}  // namespace lj
