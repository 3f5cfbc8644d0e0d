/*
 * lj_oracle.h -- CPU oracle for the Lennard-Jones cutoff hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code: it is the checker used by tests/, by the smoke check and by
 * the CPU-baseline leg of bench.py.  It restates, in plain C, the reference's
 * cell-grid algorithm (Cuda-version/cudaLJ.cu, Cuda-version/prefixsum.hpp) and
 * the OptiX variants' interaction predicates (SDK/optixMulti*), each function
 * citing the reference lines it follows.
 *
 * Parity pins (what checks each function, all in tests/test_oracle.py):
 *   orc_grid_dims / orc_cell_key   numpy float32 IEEE division + clamp, brute force
 *   orc_bin                         numpy cumsum (scan), numpy stable argsort (perm)
 *   orc_candidates(_direct)         two independent formulas + brute enumeration
 *   orc_pairs_brute / orc_pairs_cellgrid
 *                                   each other (O(N^2) vs 27-cell), exact-rational
 *                                   d^2 on small sets, symmetry
 *   orc_pairs_rows                  the matching rows of orc_pairs_cellgrid
 *   orc_forces_f64                  two-body closed forms (V(sigma)=0, F(sigma)=24eps/sigma,
 *                                   minimum at 2^(1/6) sigma), simple-cubic lattice sums,
 *                                   Newton's third law, finite-difference gradient,
 *                                   permutation and translation invariance
 *   orc_forces_f32_cellgrid         orc_forces_f64 within the FP32 error bound
 *   orc_cudalj_energy               numpy FP64 restatement of cudaLJ.cu:28-56 on tiny sets
 *   orc_pair_mult (SPHERE, TRIANGLE) hand-placed pairs whose multiplicity follows from
 *                                   the geometry (generic 1, coincident-axis 2 and 4,
 *                                   panel-overlap 2, near-coincident misses), an
 *                                   independent numpy ray caster, and the distance
 *                                   predicates it reduces to on generic pairs
 * The reference itself cannot be compiled here (it needs nvcc, cuRAND and the
 * OptiX driver), so there is no oracle/_ref build.
 */
#ifndef LJ_ORACLE_H
#define LJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Predicates (same numbering as the product's lj_predicate). */
#define ORC_PRED_AABB     0 /* 1e-16 <= d2 <= rc2      SDK/optixMultiSphereAABB/optixLJ.cu:162 */
#define ORC_PRED_CUSTOM   1 /* 1e-16 <= d2 <  rc2      SDK/optixMultiCustom/optixMultiCustomAABB.cu:59 */
#define ORC_PRED_CUDALJ   2 /* every 27-cell candidate j != i, softened, Cuda-version/cudaLJ.cu:319-337 */
#define ORC_PRED_SPHERE   3 /* ray-hit model: d2 < rc2, closest-axis ray hits  SDK/optixMultiSphere/optixLJ.cu:77-194 */
#define ORC_PRED_TRIANGLE 4 /* ray-hit model, multiplicity  SDK/optixMultiTriangle/optixLJ.cu:82-217 */

typedef struct orc_params {
    float rc, sigma, epsilon, softening2;
    float box_lo[3], box_hi[3];
    float cell_edge; /* 0 -> rc */
    int32_t predicate;
    int32_t precision; /* unused by the oracle; keeps the product's lj_params layout */
    int32_t flags;
} orc_params;

/* Grid geometry, cudaLJ.cu:525-531: G = (int)(L / cell_edge) in FP32, at least 1. */
int orc_grid_dims(const orc_params* p, int32_t dims[3], int64_t* ncells);

/* Cell key of one position, cudaLJ.cu:108-134 with origin box_lo and a clamp to [0, G-1]. */
uint32_t orc_cell_key(const orc_params* p, const int32_t dims[3], float x, float y, float z);

/* Histogram + exclusive scan + stable counting sort (cudaLJ.cu:345-383, 428-471;
 * prefixsum.hpp:160-215).  offsets has ncells+1 entries; perm[k] = original index
 * of the k-th particle in cell order.  keys (nullable) receives each particle's key. */
int orc_bin(const orc_params* p, const float* x, const float* y, const float* z, int64_t n,
            uint32_t* offsets, uint32_t* perm, uint32_t* keys);

/* Candidate count of cudaLJ.cu:255-297 from the cell offsets: closed form
 * sum_c n_c * (sum of the clipped 27 cells) - N. */
uint64_t orc_candidates(const int32_t dims[3], const uint32_t* offsets);
/* Same count, per particle, literally as cudaLJ.cu:266-293 (home n-1, plus 26 others). */
uint64_t orc_candidates_direct(const orc_params* p, const float* x, const float* y,
                               const float* z, int64_t n, const uint32_t* offsets);

/* Canonical FP32 squared distance, d2 = fma(dz,dz, fma(dy,dy, dx*dx)), d = pj - pi. */
float orc_d2(float xi, float yi, float zi, float xj, float yj, float zj);
/* Predicate on one pair from its squared distance (for CUDALJ: i != j only; for
 * SPHERE and TRIANGLE only the distance part of the pair rule, see orc_pair_mult). */
int orc_accept(int pred, float d2, float rc, int64_t i, int64_t j);
/* How many times pair (i, j) enters particle i's sums: 0 or 1 for the distance
 * predicates; for SPHERE and TRIANGLE the number of any-hit calls of the reference's
 * ray programs that add the pair (ray-hit models of SDK/optixMultiSphere/optixLJ.cu:77-194
 * and SDK/optixMultiTriangle/optixLJ.cu:82-217; see lj_oracle.c).  The pair lists
 * hold the pairs with a multiplicity above 0; forces, energies and n_pairs count each
 * pair as many times as its multiplicity. */
int orc_pair_mult(const orc_params* p, float xi, float yi, float zi, float xj, float yj, float zj, int64_t i,
                  int64_t j);

/* O(N^2) neighbour lists in CSR form.  rows ordered by original index i, each row
 * sorted ascending by j.  Returns the number of ordered pairs; writes at most cap
 * entries of cols (cols may be NULL to count only). */
uint64_t orc_pairs_brute(const orc_params* p, const float* x, const float* y, const float* z,
                         int64_t n, uint64_t* row_offsets, uint32_t* cols, uint64_t cap);
/* Same lists from the 27-cell grid (candidate enumeration + predicate). */
uint64_t orc_pairs_cellgrid(const orc_params* p, const float* x, const float* y, const float* z,
                            int64_t n, uint64_t* row_offsets, uint32_t* cols, uint64_t cap);

/* The rows of orc_pairs_cellgrid for particles subset[0..nsub) only: row r of the CSR
 * output (row_offsets[nsub+1], cols) is particle subset[r], sorted ascending.  OpenMP
 * over rows; cols is written only when the total fits cap.  Returns the total. */
uint64_t orc_pairs_rows(const orc_params* p, const float* x, const float* y, const float* z, int64_t n,
                        const int64_t* subset, int64_t nsub, uint64_t* row_offsets, uint32_t* cols, uint64_t cap,
                        int nthreads);

/* FP64 forces and energies on the FP32 inputs.  The pair set is decided by the
 * FP32 canonical predicate (so it is the product's pair set); the arithmetic of
 * each accepted pair and every sum is FP64.  Outputs in original order.
 * sums[0] = total energy (1/2 sum_i e_i), sums[1] = virial (1/2 sum_ij w_ij r_ij^2),
 * counts[0] = 27-cell candidates, counts[1] = accepted ordered pairs.
 * cond (nullable) receives sum_j |f_ij| per particle (error conditioning).
 * subset (nullable, nsub entries) restricts the computation to those particles. */
int orc_forces_f64(const orc_params* p, const float* x, const float* y, const float* z, int64_t n,
                   const int64_t* subset, int64_t nsub,
                   double* fx, double* fy, double* fz, double* e, double* cond,
                   double sums[2], uint64_t counts[2], int nthreads);

/* FP32 restatement of cudaLJ's cell-grid path (bin, scan, scatter, 27-cell z->y->x
 * loop, cudaLJ.cu:600-653) with the canonical cutoff predicate and forces added.
 * OpenMP over particles with nthreads threads.  This is the timed CPU baseline.
 * Outputs in original order; sums/counts as above. */
int orc_forces_f32_cellgrid(const orc_params* p, const float* x, const float* y, const float* z,
                            int64_t n, float* fx, float* fy, float* fz, float* e,
                            double sums[2], uint64_t counts[2], int nthreads);

/* Reference cudaLJ energy kernel semantics (cudaLJ.cu:28-56, 299-341): per particle
 * sum over every 27-cell candidate j != i of 4eps((s/r)^12-(s/r)^6), r = sqrt(d2+soft).
 * FP64 arithmetic on FP32 inputs; output in original order. */
int orc_cudalj_energy(const orc_params* p, const float* x, const float* y, const float* z,
                      int64_t n, double* e);

/* Philox4x64-10 block function (Salmon et al., SC'11). */
void orc_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
