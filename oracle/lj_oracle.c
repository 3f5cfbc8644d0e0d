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


uint64_t orc_pairs_brute(const orc_params* p, const float* x, const float* y, const float* z,
                         int64_t n, uint64_t* row_offsets, uint32_t* cols, uint64_t cap) {
    uint64_t m = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (row_offsets) row_offsets[i] = m;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            const float d2 = orc_d2(x[i], y[i], z[i], x[j], y[j], z[j]);
            if (orc_accept(p->predicate, d2, p->rc, i, j)) {
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
                        const float d2 = orc_d2(x[i], y[i], z[i], x[j], y[j], z[j]);
                        if (orc_accept(p->predicate, d2, p->rc, i, j)) {
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
                        const float d2f = orc_d2(x[i], y[i], z[i], x[j], y[j], z[j]);
                        if (!orc_accept(p->predicate, d2f, p->rc, i, j)) continue;
                        ++np_i;
                        const double dx = (double)x[j] - (double)x[i];
                        const double dy = (double)y[j] - (double)y[i];
                        const double dz = (double)z[j] - (double)z[i];
                        const double r2 = dx * dx + dy * dy + dz * dz + soft;
                        const double s2 = sig2 / r2;
                        const double s6 = s2 * s2 * s2;
                        const double w = 24.0 * eps * s6 * (2.0 * s6 - 1.0) / r2;
                        ax -= w * dx;
                        ay -= w * dy;
                        az -= w * dz;
                        ac += fabs(w) * sqrt(dx * dx + dy * dy + dz * dz);
                        ae += 4.0 * eps * s6 * (s6 - 1.0);
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
                    if (!orc_accept(p->predicate, d2, p->rc, s, t)) continue;
                    ++np_i;
                    const float r2 = d2 + soft;
                    const float ir2 = 1.0f / r2;
                    const float s2 = sig2 * ir2;
                    const float s6 = s2 * s2 * s2;
                    const float w = e24 * s6 * (2.0f * s6 - 1.0f) * ir2;
                    ax -= w * dx;
                    ay -= w * dy;
                    az -= w * dz;
                    ae += e4 * s6 * (s6 - 1.0f);
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
