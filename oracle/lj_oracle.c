/*
 * lj_oracle.c -- CPU oracle for the Lennard-Jones cutoff hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, the smoke check and the
 * CPU-baseline leg of bench.py.  The product library never links or calls it.
 * Parity pins for every function are listed in lj_oracle.h.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).  FP contraction is
 * off so that every fused multiply-add is the explicit fmaf() of the canonical
 * squared distance and nothing else.
 */
#include "lj_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Philox4x64-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11).  Multipliers and Weyl key increments from that paper. */
void orc_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
    uint64_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint64_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        unsigned __int128 p0 = (unsigned __int128)0xD2E7470EE14C6C93ull * x0;
        unsigned __int128 p1 = (unsigned __int128)0xCA5A826395121157ull * x2;
        uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
        uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
        x0 = hi1 ^ x1 ^ k0;
        x1 = lo1;
        x2 = hi0 ^ x3 ^ k1;
        x3 = lo0;
        k0 += 0x9E3779B97F4A7C15ull;
        k1 += 0xBB67AE8584CAA73Bull;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* ------------------------------------------------------------------------- */
/* Geometry: Cuda-version/cudaLJ.cu:525-531 computes GridDim = (int)(Box/cutoff)
 * in FP32 with origin 0.  The origin is generalised to box_lo; with box_lo = 0,
 * x - 0 == x exactly, so the reference rule is reproduced bit for bit. */
static float cell_edge_of(const orc_params* p) { return p->cell_edge > 0.f ? p->cell_edge : p->rc; }

int orc_grid_dims(const orc_params* p, int32_t dims[3], int64_t* ncells) {
    const float cw = cell_edge_of(p);
    if (!(cw > 0.f) || !isfinite(cw)) return -1;
    int64_t nc = 1;
    for (int d = 0; d < 3; ++d) {
        const float L = p->box_hi[d] - p->box_lo[d];
        if (!(L > 0.f) || !isfinite(L)) return -1;
        const float q = L / cw;
        if (!(q < 16777216.0f)) return -5;
        int32_t g = (int32_t)q;
        if (g < 1) g = 1;
        dims[d] = g;
        nc *= g;
    }
    if (nc > 0x7fffffffll) return -5;
    *ncells = nc;
    return 0;
}

/* PosToCellCoord (cudaLJ.cu:108-118): (int)(x / cw), IEEE FP32 division and
 * truncation, here on x - box_lo.  The clamp to [0, G-1] is new: the largest float
 * below 1.0 divided by 0.04/0.01/0.005 rounds to G, and the reference only avoids
 * that through its own sweep values (cudaLJ.cu:185,753).  fmaxf sends NaN to 0. */
static inline int32_t axis_cell(float x, float lo, float cw, int32_t g) {
    float q = (x - lo) / cw;
    q = fmaxf(q, 0.0f);
    q = fminf(q, (float)(g - 1));
    return (int32_t)q;
}

/* CellCoordToIndex (cudaLJ.cu:127-134): ((z*Gy)+y)*Gx+x, z-major. */
uint32_t orc_cell_key(const orc_params* p, const int32_t dims[3], float x, float y, float z) {
    const float cw = cell_edge_of(p);
    const int32_t cx = axis_cell(x, p->box_lo[0], cw, dims[0]);
    const int32_t cy = axis_cell(y, p->box_lo[1], cw, dims[1]);
    const int32_t cz = axis_cell(z, p->box_lo[2], cw, dims[2]);
    return (uint32_t)(((int64_t)cz * dims[1] + cy) * dims[0] + cx);
}

/* ComputeNbParticlePerCells (cudaLJ.cu:345-383) -> histogram;
 * PrefixFullV2 into prefix+1 with prefix[0]=0 (cudaLJ.cu:621-622,
 * prefixsum.hpp:160-215) -> exclusive offsets[Nc+1];
 * MoveParticlesToNewCells (cudaLJ.cu:441-471) -> scatter.  The reference's ticket
 * order inside a cell is nondeterministic; here particles keep input order inside
 * each cell (a stable counting sort), which is the order the product defines. */
int orc_bin(const orc_params* p, const float* x, const float* y, const float* z, int64_t n,
            uint32_t* offsets, uint32_t* perm, uint32_t* keys) {
    int32_t dims[3];
    int64_t nc;
    int rc = orc_grid_dims(p, dims, &nc);
    if (rc) return rc;
    uint32_t* k = keys ? keys : (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    uint32_t* fill = (uint32_t*)calloc((size_t)nc, sizeof(uint32_t));
    memset(offsets, 0, sizeof(uint32_t) * (size_t)(nc + 1));
    for (int64_t i = 0; i < n; ++i) {
        k[i] = orc_cell_key(p, dims, x[i], y[i], z[i]);
        offsets[k[i] + 1] += 1;
    }
    for (int64_t c = 0; c < nc; ++c) offsets[c + 1] += offsets[c];
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t c = k[i];
        perm[offsets[c] + fill[c]++] = (uint32_t)i;
    }
    free(fill);
    if (!keys) free(k);
    return 0;
}

/* ComputeNbInteractions (cudaLJ.cu:255-297): per particle (n_home - 1) plus the
 * counts of the other cells of the clipped 3x3x3 stencil.  Summed per cell:
 * sum_c n_c * S_c - N with S_c the clipped 27-cell sum including c. */
uint64_t orc_candidates(const int32_t dims[3], const uint32_t* offsets) {
    const int32_t gx = dims[0], gy = dims[1], gz = dims[2];
    uint64_t total = 0, n = 0;
    for (int32_t cz = 0; cz < gz; ++cz)
        for (int32_t cy = 0; cy < gy; ++cy)
            for (int32_t cx = 0; cx < gx; ++cx) {
                const int64_t c = ((int64_t)cz * gy + cy) * gx + cx;
                const uint64_t nh = offsets[c + 1] - offsets[c];
                if (!nh) continue;
                uint64_t s = 0;
                for (int32_t z = cz - 1; z <= cz + 1; ++z) {
                    if (z < 0 || z >= gz) continue;
                    for (int32_t yy = cy - 1; yy <= cy + 1; ++yy) {
                        if (yy < 0 || yy >= gy) continue;
                        const int64_t row = ((int64_t)z * gy + yy) * gx;
                        const int32_t x0 = cx > 0 ? cx - 1 : 0;
                        const int32_t x1 = cx < gx - 1 ? cx + 1 : gx - 1;
                        s += offsets[row + x1 + 1] - offsets[row + x0];
                    }
                }
                total += nh * s;
                n += nh;
            }
    return total - n;
}

uint64_t orc_candidates_direct(const orc_params* p, const float* x, const float* y,
                               const float* z, int64_t n, const uint32_t* offsets) {
    int32_t dims[3];
    int64_t nc;
    if (orc_grid_dims(p, dims, &nc)) return 0;
    const float cw = cell_edge_of(p);
    uint64_t total = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int32_t cx = axis_cell(x[i], p->box_lo[0], cw, dims[0]);
        const int32_t cy = axis_cell(y[i], p->box_lo[1], cw, dims[1]);
        const int32_t cz = axis_cell(z[i], p->box_lo[2], cw, dims[2]);
        for (int32_t ix = cx - 1; ix <= cx + 1; ++ix)
            for (int32_t iy = cy - 1; iy <= cy + 1; ++iy)
                for (int32_t iz = cz - 1; iz <= cz + 1; ++iz) {
                    if (ix < 0 || iy < 0 || iz < 0 || ix >= dims[0] || iy >= dims[1] || iz >= dims[2])
                        continue;
                    const int64_t c = ((int64_t)iz * dims[1] + iy) * dims[0] + ix;
                    const uint64_t cnt = offsets[c + 1] - offsets[c];
                    total += (ix == cx && iy == cy && iz == cz) ? cnt - 1 : cnt;
                }
    }
    return total;
}

/* ------------------------------------------------------------------------- */
/* Canonical squared distance.  The OptiX AABB intersection program computes
 * dot(offset, offset) on offset = p_i - p_j (SDK/optixMultiSphereAABB/optixLJ.cu:151-153);
 * here the association is fixed as fma(dz,dz, fma(dy,dy, dx*dx)) so that host
 * and device agree bit for bit.  (-a)^2 == a^2 exactly, so the sign of d is moot. */
float orc_d2(float xi, float yi, float zi, float xj, float yj, float zj) {
    const float dx = xj - xi, dy = yj - yi, dz = zj - zi;
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

/* The largest float <= 1e-4 (double): sqrt(d2) > 1e-4 of
 * SDK/optixMultiTriangle/optixLJ.cu:188 compares a float against a double. */
static float tri_lo(void) {
    float t = (float)1e-4;
    if ((double)t > 1e-4) t = nextafterf(t, 0.0f);
    return t;
}

int orc_accept(int pred, float d2, float rc, int64_t i, int64_t j) {
    const float rc2 = rc * rc;
    switch (pred) {
    case ORC_PRED_AABB:   return d2 <= rc2 && d2 >= 1e-16f;
    case ORC_PRED_CUSTOM: return d2 < rc2 && d2 >= 1e-16f;
    case ORC_PRED_CUDALJ: return i != j;
    case ORC_PRED_SPHERE: return d2 < rc2 && d2 > 0.0f;
    case ORC_PRED_TRIANGLE: {
        const float d = sqrtf(d2);
        return d < rc && d > tri_lo();
    }
    }
    return 0;
}


/* ------------------------------------------------------------------------- */
/* Ray-hit models of the Sphere and Triangle programs.  Both OptiX variants find
 * pairs by casting short rays at per-particle geometry; a pair is added once per
 * any-hit call that passes the program's filter, so the number of times pair (i, j)
 * enters particle i's energy (its multiplicity) follows from the geometry.  These are
 * literal restatements: every ray of particle i is intersected with every primitive of
 * particle j in FP64, and the any-hit program's filter is applied to each hit.
 * What OptiX itself decides is closed driver code; the model assumes that a ray
 * reports a sphere once (at its first surface crossing with t in [tmin, tmax]) and a
 * two-triangle panel once (the two triangles share a diagonal, and the intersector is
 * watertight without double hits).  Hits within FP rounding of a primitive's edge or of
 * tmin / tmax, and the reference's --use_fast_math square root, are not modelled. */

/* optixMultiSphere: particle j is a sphere of radius R = c*0.81649658092f + 1e-5f
 * (SDK/optixMultiSphere/optixSphere.cpp:156-157, FP32).  Particle i casts three axis
 * rays from p_i - h e_a, h = c*0.81649658092f, direction e_a, t in [0, 2h]
 * (optixLJ.cu:77-128).  The any-hit program (optixLJ.cu:150-194) adds the pair when
 * d2 < c^2 and the ray's axis is the one closest to p_j, with ties going to x, then y,
 * then z: so each ray of i decides for the pairs nearest its own axis.  d2 is the
 * canonical squared distance here (the reference's sum of squares may be contracted
 * by the compiler). */
static int sphere_mult(float c, const float pi[3], const float pj[3]) {
    const float h = c * 0.81649658092f;
    const float rad = c * 0.81649658092f + 1e-5f;
    const double tmax = 2.0 * (double)h;
    int m = 0;
    for (int a = 0; a < 3; ++a) {
        float o[3];
        for (int b = 0; b < 3; ++b) o[b] = pi[b] - h * (b == a ? 1.0f : 0.0f);
        /* |o + t e_a - p_j|^2 = R^2 */
        const double along = (double)pj[a] - (double)o[a];
        double perp2 = 0.0;
        for (int b = 0; b < 3; ++b)
            if (b != a) {
                const double d = (double)pj[b] - (double)o[b];
                perp2 += d * d;
            }
        const double disc = (double)rad * (double)rad - perp2;
        if (disc < 0.0) continue;
        const double sq = sqrt(disc), t1 = along - sq, t2 = along + sq;
        const double t = t1 >= 0.0 ? t1 : t2; /* first crossing at or after tmin = 0 */
        if (!(t >= 0.0 && t <= tmax)) continue;
        /* any-hit body */
        const float d2 = orc_d2(pi[0], pi[1], pi[2], pj[0], pj[1], pj[2]);
        if (!(d2 < c * c)) continue;
        const float ax = fabsf(pi[0] - pj[0]), ay = fabsf(pi[1] - pj[1]), az = fabsf(pi[2] - pj[2]);
        const float sx = ax * ax, sy = ay * ay, sz = az * az;
        const float dax = sy + sz, day = sx + sz, daz = sx + sy;
        const float closest = fminf(dax, fminf(day, daz));
        const int mine = (closest == dax && a == 0) || (closest == day && a == 1 && closest != dax) ||
                         (closest == daz && a == 2 && closest != dax && closest != day);
        if (mine) ++m;
    }
    return m;
}

/* Moller & Trumbore, "Fast, minimum storage ray-triangle intersection" (1997), in
 * FP64, for a ray with direction (1, 0, 0); edges and vertices count as inside.  For a
 * triangle in a plane x = const (the panels) the hit distance is taken exactly, as the
 * plane's x minus the origin's (both FP32, so the FP64 difference is exact), rather
 * than from the rounded determinant ratio: the t range test then has one answer. */
static int ray_triangle(const float o[3], const float v0[3], const float v1[3], const float v2[3], double tmin,
                        double tmax) {
    const double e1[3] = {(double)v1[0] - v0[0], (double)v1[1] - v0[1], (double)v1[2] - v0[2]};
    const double e2[3] = {(double)v2[0] - v0[0], (double)v2[1] - v0[1], (double)v2[2] - v0[2]};
    /* pvec = dir x e2 with dir = (1, 0, 0) */
    const double pv[3] = {0.0, -e2[2], e2[1]};
    const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det == 0.0) return 0;
    const double tv[3] = {(double)o[0] - v0[0], (double)o[1] - v0[1], (double)o[2] - v0[2]};
    const double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) / det;
    if (u < 0.0 || u > 1.0) return 0;
    const double qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    const double v = qv[0] / det; /* dir . qvec */
    if (v < 0.0 || u + v > 1.0) return 0;
    const double t = (v0[0] == v1[0] && v1[0] == v2[0]) ? (double)v0[0] - (double)o[0]
                                                          : (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) / det;
    return t >= tmin && t <= tmax;
}

/* optixMultiTriangle: particle j carries two square panels normal to x, at x_j - c/2
 * and x_j + c/2, each two triangles, whose corners are built in FP32 on the host
 * (SDK/optixMultiTriangle/optixTriangle.cpp:184-222): y and z span p_j +- (c/2 +
 * 1e-5).  Particle i casts four rays along +x from (x_i - c/2, y_i +- c/2, z_i +- c/2),
 * t in [FLT_EPSILON, c + c 1e-5] (optixLJ.cu:82-131).  The any-hit program
 * (optixLJ.cu:164-217) rebuilds the centre q of the hit panel from the triangle's
 * vertices, keeps the hit when 1e-4 < |q - p_i| < c, and adds it for ray 0 always and
 * for rays 1-3 when q differs from p_i in both y and z or lies on the ray's side.
 * Pairs that sit within 1e-5 of each other in y or z are hit by two rays (four if both),
 * and pairs within c 1e-5 in x by both panels, so the multiplicity can exceed 1. */
static int triangle_mult(float c, const float pi[3], const float pj[3]) {
    const float eps = 0.00001f;
    const float hc = c / 2;
    const float lo_off = -hc - eps, hi_off = hc + eps;
    float corner[8][3];
    for (int k = 0; k < 8; ++k) {
        corner[k][2] = pj[2] + ((k & 1) ? hi_off : lo_off);
        corner[k][1] = pj[1] + ((k & 2) ? hi_off : lo_off);
        corner[k][0] = pj[0] + ((k & 4) ? hc : -hc);
    }
    static const int tri[4][3] = {{0, 1, 3}, {0, 2, 3}, {4, 5, 7}, {4, 6, 7}};
    const float half = c / 2;
    const double tmin = (double)1.19209290e-07F;
    const double tmax = (double)(float)((double)(2 * half) + (double)(2 * half) * 0.00001);
    int m = 0;
    for (int r = 0; r < 4; ++r) {
        const float ycoef = (r & 1) ? -1.0f : 1.0f, zcoef = (r & 2) ? -1.0f : 1.0f;
        const float o[3] = {pi[0] - half, pi[1] + half * ycoef, pi[2] + half * zcoef};
        int seen[2] = {0, 0};
        for (int t = 0; t < 4; ++t) {
            const int panel = t / 2;
            if (seen[panel]) continue;
            const float* v0 = corner[tri[t][0]];
            const float* v1 = corner[tri[t][1]];
            const float* v2 = corner[tri[t][2]];
            if (!ray_triangle(o, v0, v1, v2, tmin, tmax)) continue;
            seen[panel] = 1;
            /* any-hit body */
            float q[3];
            q[1] = (fmaxf(v0[1], fmaxf(v1[1], v2[1])) + fminf(v0[1], fminf(v1[1], v2[1]))) / 2;
            q[2] = (fmaxf(v0[2], fmaxf(v1[2], v2[2])) + fminf(v0[2], fminf(v1[2], v2[2]))) / 2;
            q[0] = panel == 0 ? v0[0] + c / 2 : v0[0] - c / 2;
            const float d = sqrtf(orc_d2(pi[0], pi[1], pi[2], q[0], q[1], q[2]));
            if (!(d < c && (double)d > 0.0001)) continue;
            const int ok = (pi[1] != q[1] && pi[2] != q[2]) || ((pi[2] < q[2] && r == 0) || (pi[2] > q[2] && r == 2)) ||
                           ((pi[1] < q[1] && r == 0) || (pi[1] > q[1] && r == 1)) || r == 0;
            if (ok) ++m;
        }
    }
    return m;
}

int orc_pair_mult(const orc_params* p, float xi, float yi, float zi, float xj, float yj, float zj, int64_t i,
                  int64_t j) {
    const float pi[3] = {xi, yi, zi}, pj[3] = {xj, yj, zj};
    if (i == j) return 0;
    switch (p->predicate) {
    case ORC_PRED_SPHERE: return sphere_mult(p->rc, pi, pj);
    case ORC_PRED_TRIANGLE: return triangle_mult(p->rc, pi, pj);
    default: return orc_accept(p->predicate, orc_d2(xi, yi, zi, xj, yj, zj), p->rc, i, j);
    }
}

uint64_t orc_pairs_brute(const orc_params* p, const float* x, const float* y, const float* z,
                         int64_t n, uint64_t* row_offsets, uint32_t* cols, uint64_t cap) {
    uint64_t m = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (row_offsets) row_offsets[i] = m;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            if (orc_pair_mult(p, x[i], y[i], z[i], x[j], y[j], z[j], i, j) > 0) {
                if (cols && m < cap) cols[m] = (uint32_t)j;
                ++m;
            }
        }
    }
    if (row_offsets) row_offsets[n] = m;
    return m;
}

static int cmp_u32(const void* a, const void* b) {
    const uint32_t u = *(const uint32_t*)a, v = *(const uint32_t*)b;
    return (u > v) - (u < v);
}

/* Enumerates, for original particle i, every candidate j of its clipped 27-cell
 * stencil (z -> y -> x order of cudaLJ.cu:319-321), j != i. */
typedef struct grid_ctx {
    int32_t dims[3];
    int64_t nc;
    uint32_t* offsets;
    uint32_t* perm;
    uint32_t* keys;
} grid_ctx;

static int grid_build(const orc_params* p, const float* x, const float* y, const float* z,
                      int64_t n, grid_ctx* g) {
    int rc = orc_grid_dims(p, g->dims, &g->nc);
    if (rc) return rc;
    g->offsets = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(g->nc + 1));
    g->perm = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    g->keys = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    return orc_bin(p, x, y, z, n, g->offsets, g->perm, g->keys);
}

static void grid_free(grid_ctx* g) {
    free(g->offsets);
    free(g->perm);
    free(g->keys);
}

uint64_t orc_pairs_cellgrid(const orc_params* p, const float* x, const float* y, const float* z,
                            int64_t n, uint64_t* row_offsets, uint32_t* cols, uint64_t cap) {
    grid_ctx g;
    if (grid_build(p, x, y, z, n, &g)) return 0;
    const int32_t gx = g.dims[0], gy = g.dims[1], gz = g.dims[2];
    uint64_t m = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (row_offsets) row_offsets[i] = m;
        const uint64_t row_start = m;
        const uint32_t key = g.keys[i];
        const int32_t cx = (int32_t)(key % (uint32_t)gx);
        const int32_t cy = (int32_t)((key / (uint32_t)gx) % (uint32_t)gy);
        const int32_t cz = (int32_t)(key / ((uint32_t)gx * (uint32_t)gy));
        for (int32_t iz = cz - 1; iz <= cz + 1; ++iz) {
            if (iz < 0 || iz >= gz) continue;
            for (int32_t iy = cy - 1; iy <= cy + 1; ++iy) {
                if (iy < 0 || iy >= gy) continue;
                for (int32_t ix = cx - 1; ix <= cx + 1; ++ix) {
                    if (ix < 0 || ix >= gx) continue;
                    const int64_t c = ((int64_t)iz * gy + iy) * gx + ix;
                    for (uint32_t s = g.offsets[c]; s < g.offsets[c + 1]; ++s) {
                        const uint32_t j = g.perm[s];
                        if ((int64_t)j == i) continue;
                        if (orc_pair_mult(p, x[i], y[i], z[i], x[j], y[j], z[j], i, j) > 0) {
                            if (cols && m < cap) cols[m] = j;
                            ++m;
                        }
                    }
                }
            }
        }
        if (cols && m <= cap) qsort(cols + row_start, (size_t)(m - row_start), sizeof(uint32_t), cmp_u32);
    }
    if (row_offsets) row_offsets[n] = m;
    grid_free(&g);
    return m;
}

/* The rows of orc_pairs_cellgrid for a subset of particles only (row r of the
 * output is particle subset[r]), OpenMP over rows: for checking a few thousand rows
 * of a large system.  Two passes: counts, then the sorted rows. */
static uint64_t row_fill(const orc_params* p, const grid_ctx* g, const float* x, const float* y, const float* z,
                         int64_t i, uint32_t* out) {
    const int32_t gx = g->dims[0], gy = g->dims[1], gz = g->dims[2];
    const uint32_t key = g->keys[i];
    const int32_t cx = (int32_t)(key % (uint32_t)gx);
    const int32_t cy = (int32_t)((key / (uint32_t)gx) % (uint32_t)gy);
    const int32_t cz = (int32_t)(key / ((uint32_t)gx * (uint32_t)gy));
    uint64_t m = 0;
    for (int32_t iz = cz - 1; iz <= cz + 1; ++iz) {
        if (iz < 0 || iz >= gz) continue;
        for (int32_t iy = cy - 1; iy <= cy + 1; ++iy) {
            if (iy < 0 || iy >= gy) continue;
            for (int32_t ix = cx - 1; ix <= cx + 1; ++ix) {
                if (ix < 0 || ix >= gx) continue;
                const int64_t c = ((int64_t)iz * gy + iy) * gx + ix;
                for (uint32_t s = g->offsets[c]; s < g->offsets[c + 1]; ++s) {
                    const uint32_t j = g->perm[s];
                    if ((int64_t)j == i) continue;
                    if (orc_pair_mult(p, x[i], y[i], z[i], x[j], y[j], z[j], i, j) > 0) {
                        if (out) out[m] = j;
                        ++m;
                    }
                }
            }
        }
    }
    if (out) qsort(out, (size_t)m, sizeof(uint32_t), cmp_u32);
    return m;
}

uint64_t orc_pairs_rows(const orc_params* p, const float* x, const float* y, const float* z, int64_t n,
                        const int64_t* subset, int64_t nsub, uint64_t* row_offsets, uint32_t* cols, uint64_t cap,
                        int nthreads) {
    grid_ctx g;
    if (grid_build(p, x, y, z, n, &g)) return 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    uint64_t* cnt = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(nsub > 0 ? nsub : 1));
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < nsub; ++r) cnt[r] = row_fill(p, &g, x, y, z, subset[r], NULL);
    uint64_t m = 0;
    for (int64_t r = 0; r < nsub; ++r) {
        row_offsets[r] = m;
        m += cnt[r];
    }
    row_offsets[nsub] = m;
    if (cols && m <= cap) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t r = 0; r < nsub; ++r) row_fill(p, &g, x, y, z, subset[r], cols + row_offsets[r]);
    }
    free(cnt);
    grid_free(&g);
    return m;
}

/* ------------------------------------------------------------------------- */
/* Lennard-Jones pair in FP64.  V(r) = 4 eps ((s/r)^12 - (s/r)^6) as in
 * lennardJonesPotential (cudaLJ.cu:35-56, SDK/optixMultiSphereAABB/optixLJ.cu:113-124);
 * the force is its analytic gradient, F_i = 24 eps (2 s^12 - s^6) / r^2 * (r_i - r_j),
 * which the reference does not compute.  r^2 carries the softening term of
 * distance() (cudaLJ.cu:28-33) in CUDALJ mode. */
int orc_forces_f64(const orc_params* p, const float* x, const float* y, const float* z, int64_t n,
                   const int64_t* subset, int64_t nsub,
                   double* fx, double* fy, double* fz, double* e, double* cond,
                   double sums[2], uint64_t counts[2], int nthreads) {
    grid_ctx g;
    int rc = grid_build(p, x, y, z, n, &g);
    if (rc) return rc;
    const double sig2 = (double)p->sigma * (double)p->sigma;
    const double eps = (double)p->epsilon;
    const double soft = p->predicate == ORC_PRED_CUDALJ ? (double)p->softening2 : 0.0;
    const int32_t gx = g.dims[0], gy = g.dims[1], gz = g.dims[2];
    const int64_t m = subset ? nsub : n;
    uint64_t ncand = 0, npair = 0;
    double etot = 0.0, vir = 0.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    /* Per-row results are independent; the final sums are taken serially in row
     * order so the FP64 totals do not depend on the thread count. */
    double* vrow = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    uint64_t* crow = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(m > 0 ? m : 1) * 2);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < m; ++r) {
        const int64_t i = subset ? subset[r] : r;
        const uint32_t key = g.keys[i];
        const int32_t cx = (int32_t)(key % (uint32_t)gx);
        const int32_t cy = (int32_t)((key / (uint32_t)gx) % (uint32_t)gy);
        const int32_t cz = (int32_t)(key / ((uint32_t)gx * (uint32_t)gy));
        double ax = 0, ay = 0, az = 0, ae = 0, ac = 0, av = 0;
        uint64_t nc_i = 0, np_i = 0;
        for (int32_t iz = cz - 1; iz <= cz + 1; ++iz) {
            if (iz < 0 || iz >= gz) continue;
            for (int32_t iy = cy - 1; iy <= cy + 1; ++iy) {
                if (iy < 0 || iy >= gy) continue;
                for (int32_t ix = cx - 1; ix <= cx + 1; ++ix) {
                    if (ix < 0 || ix >= gx) continue;
                    const int64_t c = ((int64_t)iz * gy + iy) * gx + ix;
                    for (uint32_t s = g.offsets[c]; s < g.offsets[c + 1]; ++s) {
                        const uint32_t j = g.perm[s];
                        if ((int64_t)j == i) continue;
                        ++nc_i;
                        const int mult = orc_pair_mult(p, x[i], y[i], z[i], x[j], y[j], z[j], i, j);
                        if (!mult) continue;
                        np_i += (uint64_t)mult;
                        const double dx = (double)x[j] - (double)x[i];
                        const double dy = (double)y[j] - (double)y[i];
                        const double dz = (double)z[j] - (double)z[i];
                        const double r2 = dx * dx + dy * dy + dz * dz + soft;
                        const double s2 = sig2 / r2;
                        const double s6 = s2 * s2 * s2;
                        const double w = mult * (24.0 * eps * s6 * (2.0 * s6 - 1.0) / r2);
                        ax -= w * dx;
                        ay -= w * dy;
                        az -= w * dz;
                        ac += fabs(w) * sqrt(dx * dx + dy * dy + dz * dz);
                        ae += mult * (4.0 * eps * s6 * (s6 - 1.0));
                        av += w * r2;
                    }
                }
            }
        }
        if (fx) fx[r] = ax;
        if (fy) fy[r] = ay;
        if (fz) fz[r] = az;
        if (e) e[r] = ae;
        if (cond) cond[r] = ac;
        vrow[r] = av;
        crow[2 * r] = nc_i;
        crow[2 * r + 1] = np_i;
    }
    for (int64_t r = 0; r < m; ++r) {
        double er = 0.0;
        if (e) er = e[r];
        etot += er;
        vir += vrow[r];
        ncand += crow[2 * r];
        npair += crow[2 * r + 1];
    }
    free(vrow);
    free(crow);
    sums[0] = 0.5 * etot;
    sums[1] = 0.5 * vir;
    counts[0] = ncand;
    counts[1] = npair;
    grid_free(&g);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* The CPU baseline: cudaLJ's loop body (cudaLJ.cu:600-653) on the host in FP32.
 *   bin    -> ComputeNbParticlePerCells + PrefixFullV2 + InitNewCells (:614-628)
 *   sort   -> MoveParticlesToNewCells into sorted SoA copies (:631-634)
 *   force  -> ComputeParticleInterationsParPartNoLoop (:299-341): per sorted
 *             particle, clipped 3x3x3 cells in z->y->x order, here with the
 *             canonical cutoff predicate and the force added.
 * The candidate count of ComputeNbInteractions (:637-646) comes from the offsets. */
int orc_forces_f32_cellgrid(const orc_params* p, const float* x, const float* y, const float* z,
                            int64_t n, float* fx, float* fy, float* fz, float* e,
                            double sums[2], uint64_t counts[2], int nthreads) {
    grid_ctx g;
    int rc = grid_build(p, x, y, z, n, &g);
    if (rc) return rc;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const size_t nn = (size_t)(n > 0 ? n : 1);
    float* sx = (float*)malloc(sizeof(float) * nn);
    float* sy = (float*)malloc(sizeof(float) * nn);
    float* sz = (float*)malloc(sizeof(float) * nn);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < n; ++s) {
        sx[s] = x[g.perm[s]];
        sy[s] = y[g.perm[s]];
        sz[s] = z[g.perm[s]];
    }
    const float sig2 = p->sigma * p->sigma;
    const float e4 = 4.0f * p->epsilon, e24 = 24.0f * p->epsilon;
    const float soft = p->predicate == ORC_PRED_CUDALJ ? p->softening2 : 0.0f;
    const int32_t gx = g.dims[0], gy = g.dims[1], gz = g.dims[2];
    double etot = 0.0, vir = 0.0;
    uint64_t npair = 0;
#pragma omp parallel for schedule(static) reduction(+ : etot, vir, npair)
    for (int64_t s = 0; s < n; ++s) {
        const uint32_t i = g.perm[s];
        const uint32_t key = g.keys[i];
        const int32_t cx = (int32_t)(key % (uint32_t)gx);
        const int32_t cy = (int32_t)((key / (uint32_t)gx) % (uint32_t)gy);
        const int32_t cz = (int32_t)(key / ((uint32_t)gx * (uint32_t)gy));
        const float xi = sx[s], yi = sy[s], zi = sz[s];
        float ax = 0.f, ay = 0.f, az = 0.f, ae = 0.f, av = 0.f;
        uint64_t np_i = 0;
        for (int32_t iz = cz - 1; iz <= cz + 1; ++iz) {
            if (iz < 0 || iz >= gz) continue;
            for (int32_t iy = cy - 1; iy <= cy + 1; ++iy) {
                if (iy < 0 || iy >= gy) continue;
                /* the three x-cells of one (y,z) row are contiguous in cell order */
                const int64_t row = ((int64_t)iz * gy + iy) * gx;
                const int32_t x0 = cx > 0 ? cx - 1 : 0;
                const int32_t x1 = cx < gx - 1 ? cx + 1 : gx - 1;
                const uint32_t b = g.offsets[row + x0], en = g.offsets[row + x1 + 1];
                for (uint32_t t = b; t < en; ++t) {
                    if ((int64_t)t == s) continue;
                    const float dx = sx[t] - xi, dy = sy[t] - yi, dz = sz[t] - zi;
                    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    int mult;
                    if (p->predicate == ORC_PRED_SPHERE || p->predicate == ORC_PRED_TRIANGLE)
                        mult = orc_pair_mult(p, xi, yi, zi, sx[t], sy[t], sz[t], s, t);
                    else
                        mult = orc_accept(p->predicate, d2, p->rc, s, t);
                    if (!mult) continue;
                    np_i += (uint64_t)mult;
                    const float r2 = d2 + soft;
                    const float ir2 = 1.0f / r2;
                    const float s2 = sig2 * ir2;
                    const float s6 = s2 * s2 * s2;
                    const float w = (float)mult * (e24 * s6 * (2.0f * s6 - 1.0f) * ir2);
                    ax -= w * dx;
                    ay -= w * dy;
                    az -= w * dz;
                    ae += (float)mult * (e4 * s6 * (s6 - 1.0f));
                    av += w * r2;
                }
            }
        }
        if (fx) fx[i] = ax;
        if (fy) fy[i] = ay;
        if (fz) fz[i] = az;
        if (e) e[i] = ae;
        etot += ae;
        vir += av;
        npair += np_i;
    }
    sums[0] = 0.5 * etot;
    sums[1] = 0.5 * vir;
    counts[0] = orc_candidates(g.dims, g.offsets);
    counts[1] = npair;
    free(sx);
    free(sy);
    free(sz);
    grid_free(&g);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* ComputeParticleInterationsParPartNoLoop (cudaLJ.cu:299-341): no cutoff test, every
 * candidate of the clipped 27 cells except the particle itself, with
 * distance() = sqrt(d2 + 1e-5) (cudaLJ.cu:28-33); FP64 arithmetic. */
int orc_cudalj_energy(const orc_params* p, const float* x, const float* y, const float* z,
                      int64_t n, double* e) {
    grid_ctx g;
    int rc = grid_build(p, x, y, z, n, &g);
    if (rc) return rc;
    const double sig = (double)p->sigma, eps = (double)p->epsilon, soft = (double)p->softening2;
    const int32_t gx = g.dims[0], gy = g.dims[1], gz = g.dims[2];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t key = g.keys[i];
        const int32_t cx = (int32_t)(key % (uint32_t)gx);
        const int32_t cy = (int32_t)((key / (uint32_t)gx) % (uint32_t)gy);
        const int32_t cz = (int32_t)(key / ((uint32_t)gx * (uint32_t)gy));
        double acc = 0.0;
        for (int32_t iz = cz - 1; iz <= cz + 1; ++iz) {
            if (iz < 0 || iz >= gz) continue;
            for (int32_t iy = cy - 1; iy <= cy + 1; ++iy) {
                if (iy < 0 || iy >= gy) continue;
                for (int32_t ix = cx - 1; ix <= cx + 1; ++ix) {
                    if (ix < 0 || ix >= gx) continue;
                    const int64_t c = ((int64_t)iz * gy + iy) * gx + ix;
                    for (uint32_t s = g.offsets[c]; s < g.offsets[c + 1]; ++s) {
                        const uint32_t j = g.perm[s];
                        if ((int64_t)j == i) continue;
                        const double dx = (double)x[j] - (double)x[i];
                        const double dy = (double)y[j] - (double)y[i];
                        const double dz = (double)z[j] - (double)z[i];
                        const double r = sqrt(dx * dx + dy * dy + dz * dz + soft);
                        const double sr = sig / r;
                        const double r6 = (sr * sr) * (sr * sr) * (sr * sr);
                        acc += 4.0 * eps * (r6 * r6 - r6);
                    }
                }
            }
        }
        e[i] = acc;
    }
    grid_free(&g);
    return 0;
}
