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
    // SPHERE ray model: ray half-length h = rc*0.81649658092f and sphere radius h+1e-5f
    float sph_h, sph_rad;
    // TRIANGLE ray model: rc/2, the panel's y/z half-extent rc/2+1e-5f (both in FP32),
    // and the ray's t range [FLT_EPSILON, rc + rc*1e-5]
    float tri_hc, tri_ext;
    double tri_tmin, tri_tmax;
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

// Ray-hit models of the two OptiX programs that find pairs with rays at particle
// geometry.  A pair enters particle i's sums once per any-hit call of the reference
// that adds it, so these return a multiplicity; every other predicate gives 0 or 1.
//
// SPHERE (SDK/optixMultiSphere/optixLJ.cu:77-194, optixSphere.cpp:156-157): particle
// j is a sphere of radius h + 1e-5 (h = rc*0.81649658092f), particle i casts one ray
// per axis from p_i - h e_a over t in [0, 2h], and the any-hit program adds the pair
// for the ray whose axis is closest to p_j (ties: x, y, z) when d2 < rc^2.  That ray
// reports the sphere when its first surface crossing at t >= 0 is within 2h: every
// pair with d2 < rc^2 except near-coincident ones, whose sphere contains the whole
// ray segment.  Evaluated for the closest axis only, in FP64 on the FP32 origin.
//
// TRIANGLE (SDK/optixMultiTriangle/optixLJ.cu:82-217, optixTriangle.cpp:184-222):
// particle j carries panels normal to x at x_j -+ rc/2 spanning y_j, z_j -+ (rc/2 +
// 1e-5); particle i casts four +x rays from (x_i - rc/2, y_i +- rc/2, z_i +- rc/2)
// over t in [FLT_EPSILON, rc (1 + 1e-5)].  A ray hits a panel when it lies within
// the panel's y and z extents and the panel's plane within its t range; the any-hit
// program rebuilds the panel centre q from its corners and adds the hit when
// 1e-4 < |q - p_i| < rc and ray 0, or q off p_i in y and z, or q on the ray's side.
template <int PRED>
__device__ inline int pair_mult(const PairConst& c, const float4& pi, const float4& pj, float d2) {
    if constexpr (PRED == LJ_PRED_SPHERE) {
        if (!(d2 < c.rc2)) return 0;
        const float ax = fabsf(pi.x - pj.x), ay = fabsf(pi.y - pj.y), az = fabsf(pi.z - pj.z);
        const float sx = ax * ax, sy = ay * ay, sz = az * az;
        const float dax = sy + sz, day = sx + sz, daz = sx + sy;
        // the closest axis, first minimum in x, y, z order
        const int a = (dax <= day && dax <= daz) ? 0 : (day <= daz ? 1 : 2);
        const float pa = a == 0 ? pi.x : (a == 1 ? pi.y : pi.z);
        const float qa = a == 0 ? pj.x : (a == 1 ? pj.y : pj.z);
        const float o = pa - c.sph_h;  // the ray origin on that axis, rounded as the reference
        const double along = (double)qa - (double)o;
        const double u = a == 0 ? (double)pj.y - (double)pi.y : (double)pj.x - (double)pi.x;
        const double v = a == 2 ? (double)pj.y - (double)pi.y : (double)pj.z - (double)pi.z;
        const double rad = (double)c.sph_rad;
        const double disc = rad * rad - (u * u + v * v);  // no contraction in this build
        if (disc < 0.0) return 0;
        const double s = __builtin_sqrt(disc), t1 = along - s;
        const double t = t1 >= 0.0 ? t1 : along + s;  // the first crossing at t >= 0
        return (t >= 0.0 && t <= 2.0 * (double)c.sph_h) ? 1 : 0;
    } else if constexpr (PRED == LJ_PRED_TRIANGLE) {
        const float lo_y = pj.y - c.tri_ext, hi_y = pj.y + c.tri_ext;
        const float lo_z = pj.z - c.tri_ext, hi_z = pj.z + c.tri_ext;
        const float yp = pi.y + c.tri_hc, ym = pi.y - c.tri_hc, zp = pi.z + c.tri_hc, zm = pi.z - c.tri_hc;
        // ray r has y offset + (r & 1 clear) or - and z offset + (r & 2 clear) or -
        const bool in_yp = lo_y <= yp && yp <= hi_y, in_ym = lo_y <= ym && ym <= hi_y;
        const bool in_zp = lo_z <= zp && zp <= hi_z, in_zm = lo_z <= zm && zm <= hi_z;
        if (!((in_yp || in_ym) && (in_zp || in_zm))) return 0;
        const float qy = (lo_y + hi_y) * 0.5f, qz = (lo_z + hi_z) * 0.5f;
        const bool off_yz = pi.y != qy && pi.z != qz;
        const bool ray_ok[4] = {true, off_yz || pi.y > qy, off_yz || pi.z > qz, off_yz};
        const bool ray_in[4] = {in_yp && in_zp, in_ym && in_zp, in_yp && in_zm, in_ym && in_zm};
        const double ox = (double)(pi.x - c.tri_hc);
        int m = 0;
#pragma unroll
        for (int panel = 0; panel < 2; ++panel) {
            const float xp = panel == 0 ? pj.x - c.tri_hc : pj.x + c.tri_hc;
            const double t = (double)xp - ox;
            if (!(t >= c.tri_tmin && t <= c.tri_tmax)) continue;
            const float qx = panel == 0 ? xp + c.tri_hc : xp - c.tri_hc;
            const float d = __builtin_sqrtf(dist2(qx - pi.x, qy - pi.y, qz - pi.z));
            if (!(d < c.rc && d > c.tri_lo)) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) m += (ray_in[r] && ray_ok[r]) ? 1 : 0;
        }
        return m;
    } else {
        return accept<PRED>(c, d2) ? 1 : 0;
    }
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
