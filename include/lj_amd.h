/*
 * lj_amd.h -- C ABI of the MI355X (gfx950) Lennard-Jones cutoff engine.
 *
 * The reference has no library interface: its hot path sits behind two host
 * drivers, executeSimulation<float>() (Cuda-version/cudaLJ.cu:515-677) and
 * core() (SDK/optixMultiSphereAABB/optixSphere.cpp:82-584), whose kernels take
 * a ParticlesContainer SoA (cudaLJ.cu:58-86) plus CellDescriptor cells
 * (cudaLJ.cu:88-94), or ParamsLJ (SDK/optixMultiSphereAABB/optixSphere.h:35-42)
 * for optixLaunch.  Each entry point below names the reference code it replaces.
 *
 * Conventions
 *  - Every call returns an int status: LJ_OK (0) or a negative LJ_ERR_* code;
 *    lj_last_error() gives the message.  Nothing exits or throws across the ABI
 *    (the reference calls exit(1) in CUDA_ASSERT, cudaLJ.cu:17-25).
 *  - Device calls take device pointers and a hipStream_t passed as void*
 *    (NULL = the default stream).  They are stream-ordered and asynchronous:
 *    outputs are valid after the stream is synchronised.
 *  - Positions are FP32 SoA x[n], y[n], z[n] in the caller's order; every
 *    per-particle output is written in that same order (the reference leaves
 *    results in cell order next to an index array, cudaLJ.cu:191,469).
 *  - A context owns a grow-only device workspace; once it has grown to a
 *    problem size, later calls of that size allocate nothing.  A context is not
 *    thread-safe: use one per host thread.
 */
#ifndef LJ_AMD_H
#define LJ_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LJ_OK 0
#define LJ_ERR_INVALID_ARG (-1)
#define LJ_ERR_OOM (-2)
#define LJ_ERR_CAPACITY (-3)
#define LJ_ERR_HIP (-4)
#define LJ_ERR_OVERFLOW (-5)
#define LJ_ERR_NO_DEVICE (-6)

/* Pair predicates.  LJ_PRED_AABB is the canonical one: d = p_j - p_i in FP32,
 * d2 = fma(dz,dz, fma(dy,dy, dx*dx)), accept iff 1e-16 <= d2 <= rc*rc.
 *   AABB     SDK/optixMultiSphereAABB/optixLJ.cu:162   1e-16 <= d2 <= rc2
 *   CUSTOM   SDK/optixMultiCustom/optixMultiCustomAABB.cu:59   1e-16 <= d2 < rc2
 *   CUDALJ   Cuda-version/cudaLJ.cu:319-337  every 27-cell candidate j != i, no cutoff,
 *            r2 = d2 + softening2 (cudaLJ.cu:28-33)
 *   SPHERE   ray-hit model of SDK/optixMultiSphere/optixLJ.cu:77-194 (optixSphere.cpp:156-157):
 *            d2 < rc2 and particle i's axis ray closest to p_j (ties x, y, z), cast from
 *            p_i - h e_a over [0, 2h], h = rc*0.81649658092f, reaches the sphere of radius
 *            h + 1e-5f around p_j (near-coincident pairs are missed)
 *   TRIANGLE ray-hit model of SDK/optixMultiTriangle/optixLJ.cu:82-217 (optixTriangle.cpp:
 *            184-222): four +x rays of particle i and two panels of particle j; each
 *            ray-panel hit with 1e-4 < |q - p_i| < rc (q the panel's rebuilt centre) that
 *            passes the ray-index rule adds the pair, so a pair can count 2 or 4 times
 * For SPHERE and TRIANGLE a pair's multiplicity (the reference's number of any-hit
 * calls that add it) weights its force and energy terms and counts in n_pairs; the
 * neighbour-list calls list each partner with a multiplicity above 0 once. */
#define LJ_PRED_AABB 0
#define LJ_PRED_CUSTOM 1
#define LJ_PRED_CUDALJ 2
#define LJ_PRED_SPHERE 3
#define LJ_PRED_TRIANGLE 4

/* lj_params.precision: how an accepted pair is evaluated.  The predicate is always
 * the FP32 canonical one, so the pair set does not depend on the precision.
 *   MIXED      (default) the pair chain r2 -> s6 -> force weight in FP64 on the exact
 *              FP32 coordinate differences, FP64 per-particle sums; on CDNA an FP64
 *              FMA issues at the rate of an unpacked FP32 FMA
 *   FP32       FP32 pair chain, plain FP32 sums (the reference's arithmetic)
 *   FP32_COMP  FP32 pair chain, FP32 sums with a TwoSum error term (compensated) */
#define LJ_PREC_MIXED 0
#define LJ_PREC_FP32 1
#define LJ_PREC_FP32_COMP 2

/* lj_params.flags */
#define LJ_FLAG_NO_VIRIAL 1u /* skip the virial sum */
/* Testing aid for the LBVH calls: caps every tree walk at one step, so that the walk's
 * step guard trips and the calls take their error path (lj_lbvh_compute marks its
 * results, lj_lbvh_pairs returns LJ_ERR_OVERFLOW).  No effect on the other calls. */
#define LJ_FLAG_WALK_GUARD_TEST 2u



/* Generators for lj_generate. */
#define LJ_GEN_UNIFORM 0  /* gparams {lo, hi}: i.i.d. U[lo,hi)^3 */
#define LJ_GEN_SURFACE 1  /* gparams {cx, cy, cz, radius, clamp_hi}: uniform on a sphere */
#define LJ_GEN_CLUSTERS 2 /* gparams {nclusters, std, cmin, cmax}: isotropic Gaussian clusters */

/* Run parameters.  Replaces the compile-time sigma = epsilon = 1 of cudaLJ.cu:332
 * and SDK/optixMultiSphereAABB/optixLJ.cu:164-165, the cutoff argument of
 * executeSimulation/core, and the implicit origin-0 box of cudaLJ.cu:525-531. */
typedef struct lj_params {
    float rc;          /* cutoff radius */
    float sigma;       /* LJ sigma */
    float epsilon;     /* LJ epsilon */
    float softening2;  /* added to d2 in LJ_PRED_CUDALJ mode (reference: 1e-5) */
    float box_lo[3];   /* grid origin */
    float box_hi[3];   /* grid extent; G = (int)((hi-lo)/cell_edge) per axis */
    float cell_edge;   /* 0 -> rc; must be >= rc for cutoff predicates */
    int32_t predicate; /* LJ_PRED_* */
    int32_t precision; /* LJ_PREC_* */
    uint32_t flags;    /* LJ_FLAG_* */
    /* Kernel shape (0 = the library's default; lj_autotune picks measured values).
     * The shape changes speed only: pair counts are identical for every shape, and
     * forces and energies agree to the rounding of the final FP32 outputs. */
    int16_t refine;    /* force-grid cells per cell_edge along y and z: 0, 1, 2 or 4;
                          the candidate count keeps the cell_edge grid of the reference */
    int16_t refine_x;  /* force-grid cells per cell_edge along x: 1, 2, 4, 8 or 16, not
                          below refine; 0 = refine when refine is set, else the default */
    int16_t block;     /* particles per workgroup: 0, 64, 128 or 256 */
    int16_t depth;     /* LDS queue entries per particle: 0, 8, 16 or 32; 1 = no queue */
} lj_params;

/* Scalar results, written to DEVICE memory by the compute calls. */
typedef struct lj_result {
    double e_total;                  /* 1/2 sum_i e_i, FP64 reduction */
    double virial;                   /* 1/2 sum_i sum_j w_ij r_ij^2, FP64 reduction */
    unsigned long long n_candidates; /* 27-cell candidates, ComputeNbInteractions semantics */
    unsigned long long n_pairs;      /* accepted ordered pairs (SPHERE, TRIANGLE: summed multiplicities) */
} lj_result;

typedef struct lj_context lj_context;

/* Library version string. */
const char* lj_version(void);

/* Creates a context on HIP device `device`.  Replaces the per-run set-up of
 * executeSimulation (allocations and stream, cudaLJ.cu:554-580) and the
 * cudaFree(0)/context creation of core() (optixSphere.cpp:96-107). */
int lj_context_create(lj_context** out, int device);
/* Frees the context and its workspace (deletePtrs, cudaLJ.cu:485-491). */
int lj_context_destroy(lj_context* ctx);
/* Message of the last failing call on ctx (NULL ctx: the last failing
 * lj_context_create in this thread). */
const char* lj_last_error(const lj_context* ctx);

/* Host: the context's workspace generation, which changes whenever a call grows
 * (reallocates) any workspace buffer.  No reference counterpart (the reference
 * allocates and frees its buffers on every run, cudaLJ.cu:554-580, 474-497): a
 * captured graph holds workspace addresses, so a caller records the generation after
 * capturing and replays only while it is unchanged.  Testing aid: the environment
 * variable LJ_TEST_FAIL_ALLOC_AFTER=k, read by lj_context_create, lets k more device
 * allocations succeed in the process and makes every later one fail with LJ_ERR_OOM. */
int lj_workspace_generation(lj_context* ctx, uint64_t* generation);

/* Host: grid dimensions and cell count for params (cudaLJ.cu:525-531). */
int lj_grid_dims(const lj_params* p, int32_t dims[3], int64_t* n_cells);

/* Host: seeded Philox4x64-10 positions into host arrays x,y,z[n].  Particle i of
 * stream s uses counter (i, s, 0, 0) and key (seed, tag).  Replaces InitParticles /
 * InitParticlesSphere (cudaLJ.cu:174-220) and the drand48 generators of
 * SDK/optixMultiSphereAABB/optixSphere.cpp:120-143. */
int lj_generate(int kind, uint64_t seed, uint64_t stream, int64_t n, const double* gparams,
                float* x, float* y, float* z);

/* Device: bins particles into the cell grid.  Replaces ComputeNbParticlePerCells +
 * PrefixFullV2 + InitNewCells + MoveParticlesToNewCells + ComputeNbInteractions
 * (cudaLJ.cu:613-646; prefixsum.hpp:160-215).
 *   cell_offsets[n_cells+1]  exclusive offsets (cell c holds sorted slots
 *                            [off[c], off[c+1]); nullable
 *   perm[n]                  original index of each sorted slot, stable; nullable
 *   n_candidates             device u64, 27-cell candidate count; nullable */
int lj_bin(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
           const lj_params* p, uint32_t* cell_offsets, uint32_t* perm,
           unsigned long long* n_candidates, void* stream);

/* Device: reorders positions into cell order (the optional -us pre-sort of
 * SDK/optixMultiSphereAABB/sort.cu:28-41).  xo,yo,zo[n] receive the sorted
 * positions, perm[n] (nullable) the original indices. */
int lj_reorder(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
               const lj_params* p, float* xo, float* yo, float* zo, uint32_t* perm, void* stream);

/* Device: forces and energies on the cell grid.  Replaces the loop body of
 * executeSimulation (cudaLJ.cu:613-653) and ComputeParticleInterationsParPartNoLoop
 * (cudaLJ.cu:299-341), with the cutoff predicate of p->predicate and the force
 * added.  fx,fy,fz[n] required; e[n] (per-particle energy, full-list sum) and
 * result (device lj_result) nullable. */
int lj_compute(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
               const lj_params* p, float* fx, float* fy, float* fz, float* e,
               lj_result* result, void* stream);

/* Device: lj_compute with the particles left in cell order, the data flow of the
 * reference's loop, which moves its particles into cell order every step
 * (MoveParticlesToNewCells, cudaLJ.cu:441-471; InitNewCells, :428-439) before the
 * energy kernel writes per-particle values in that order (cudaLJ.cu:299-341, 649-653).
 * On return x,y,z[n] hold the positions in the stable cell order of lj_bin (read first,
 * then overwritten), fx,fy,fz[n] and e[n] (nullable) the outputs of those particles in
 * the same order, and perm[n] (nullable) the index each particle had on input.  The
 * outputs are bit for bit those of lj_compute on the same input, permuted by perm; all
 * writes are consecutive, where lj_compute's go to the shuffled input order.  A caller
 * that keeps its particle state in this order also feeds the next step's sort an
 * almost sorted input.  Outputs must not alias the positions. */
int lj_compute_sorted(lj_context* ctx, float* x, float* y, float* z, int64_t n, const lj_params* p,
                      float* fx, float* fy, float* fz, float* e, uint32_t* perm, lj_result* result,
                      void* stream);

/* Device: lj_compute without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate during a call (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels and results
 * as lj_compute, bit for bit; the workspace is a fixed array of
 * LJ_NOALLOC_SCRATCH_BYTES in device memory, one per GPU, held until the process ends
 * and shared by every call on that GPU, so calls that overlap on one GPU (other
 * streams, threads or contexts) race and must be serialised by the caller.  Inputs of
 * more than LJ_NOALLOC_MAX_PARTICLES particles or a force grid (the kernel shape's
 * refined grid) of more than LJ_NOALLOC_MAX_CELLS cells return LJ_ERR_CAPACITY.
 * Profiling (lj_profile_enable) does not record this call. */
#define LJ_NOALLOC_MAX_PARTICLES (1 << 20)
#define LJ_NOALLOC_MAX_CELLS (1 << 21)
#define LJ_NOALLOC_SCRATCH_BYTES (64u << 20)
int lj_compute_noalloc(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
                       const lj_params* p, float* fx, float* fy, float* fz, float* e,
                       lj_result* result, void* stream);

/* Device: neighbour lists from the cell grid, CSR by original index.
 *   row_offsets[n+1] (device u64), cols[capacity] (device u32, original indices,
 *   each row sorted ascending).  *needed (host) receives the pair count; when it
 *   exceeds capacity nothing is written to cols and LJ_ERR_CAPACITY is returned.
 * This call synchronises the stream (the count is needed on the host). */
int lj_pairs(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
             const lj_params* p, unsigned long long* row_offsets, uint32_t* cols,
             unsigned long long capacity, unsigned long long* needed, void* stream);

/* Device: the same forces/energies through a linear BVH (Karras 2012) over the
 * per-particle boxes [p - rc, p + rc], queried with each particle's position --
 * the software counterpart of optixAccelBuild + optixLaunch with the
 * __intersection__sphere program (SDK/optixMultiSphereAABB/optixSphere.cpp:204-289,
 * 511-520; optixLJ.cu:49-88,134-176).  Same outputs as lj_compute except
 * n_candidates (leaf boxes tested).  Every walk is guarded by a step limit (2n steps,
 * which a well-formed tree never needs); if one trips, the call still returns LJ_OK
 * (it does not synchronise) but marks result: n_pairs = UINT64_MAX and NaN e_total
 * and virial, so reading the result shows the failure. */
int lj_lbvh_compute(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
                    const lj_params* p, float* fx, float* fy, float* fz, float* e,
                    lj_result* result, void* stream);

/* Device: neighbour lists through the LBVH, same contract as lj_pairs. */
int lj_lbvh_pairs(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
                  const lj_params* p, unsigned long long* row_offsets, uint32_t* cols,
                  unsigned long long capacity, unsigned long long* needed, void* stream);

/* Device: forces and energies over a stored neighbour list (CSR by original index, as
 * lj_pairs / lj_lbvh_pairs return it; entries >= n are ignored), with p's predicate
 * applied to every listed pair.  Built at rc + skin, the list holds every pair within
 * rc as long as no particle has moved more than skin / 2 since the build
 * (lj_max_displacement), and the call then finds exactly lj_compute's pairs.  Outputs
 * as lj_compute; result->n_candidates = listed pairs.  No reference counterpart: the
 * reference regenerates its particles every loop (Cuda-version/cudaLJ.cu:602-611). */
int lj_compute_list(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
                    const lj_params* p, const unsigned long long* row_offsets, const uint32_t* cols,
                    float* fx, float* fy, float* fz, float* e, lj_result* result, void* stream);

/* Device: the largest squared displacement |p - p0|^2 over n particles, into *max2
 * (one device float; NaN displacements read as infinite). */
int lj_max_displacement(lj_context* ctx, const float* x, const float* y, const float* z,
                        const float* x0, const float* y0, const float* z0, int64_t n, float* max2,
                        void* stream);

/* One system of a batch: particles [first, first+n) of the concatenated arrays. */
typedef struct lj_system {
    int64_t first;
    int64_t n;
    lj_params params;
} lj_system;

/* Device: forces/energies of many independent systems in one pipeline of
 * launches over all of them.  sys is a HOST array of nsys descriptors; results is
 * a device array of nsys lj_result.  No reference counterpart: the reference runs
 * one system per executeSimulation/core call. */
int lj_compute_batched(lj_context* ctx, const lj_system* sys, int32_t nsys, const float* x,
                       const float* y, const float* z, float* fx, float* fy, float* fz, float* e,
                       lj_result* results, void* stream);

/* Phase timing with hipEvents recorded on each call's stream.  The reference times
 * its phases with the host wall clock (SpTimer, Cuda-version/SpTimer.hpp:10-65,
 * around init and compute in cudaLJ.cu:613-659, partly without a device sync);
 * events measure the device time of each phase exactly.
 *   LJ_PHASE_BIN    cell keys, histogram, scan, radix sort, gather (LBVH: build)
 *   LJ_PHASE_FORCE  the force/energy kernel (LBVH: traversal)
 *   LJ_PHASE_OTHER  candidate count and FP64 reductions
 * lj_profile_enable resets the sums; lj_profile_read waits for the recorded events
 * and returns the summed milliseconds and the number of timed launches per phase.
 * Leave profiling off while a stream is being captured into a graph. */
#define LJ_PHASE_BIN 0
#define LJ_PHASE_FORCE 1
#define LJ_PHASE_OTHER 2
#define LJ_NUM_PHASES 3
int lj_profile_enable(lj_context* ctx, int enable);
int lj_profile_read(lj_context* ctx, double ms[3], unsigned long long calls[3]);

/* Autotuner.  Reference: none -- the reference fixes its launch shape (one thread
 * per particle in blocks of DefaultNbThreads = 128, cudaLJ.cu:519-520, 647-650) and
 * its cell edge (= cutoff, cudaLJ.cu:525-531).  lj_autotune times lj_compute on the
 * given device positions, `reps` calls per shape after one untimed call: first every
 * valid grid (refine, refine_x) with block 256 and no queue, then every (block, depth)
 * on the fastest grid.  It writes the fastest shape into p->refine, p->refine_x,
 * p->block and p->depth (no other field changes).  Outputs of the timed calls go to the context's
 * workspace.  It synchronises `stream` and records hipEvents, so it cannot be
 * captured in a graph: tune once per workload, then pass the tuned params. */
typedef struct lj_tuning {
    int32_t refine, refine_x;     /* chosen grid (also written to the params) */
    int32_t block, depth;         /* chosen workgroup size and queue depth */
    int32_t n_tried;              /* shapes timed */
    float ms_best;                /* mean time of one call with the chosen shape */
    float ms_default;             /* mean time of one call with the default shape */
    int32_t reserved;
} lj_tuning;
int lj_autotune(lj_context* ctx, const float* x, const float* y, const float* z, int64_t n,
                lj_params* p, int32_t reps, lj_tuning* out, void* stream);
/* The same for a batch (lj_compute_batched): one shape for the whole batch, written
 * into every sys[s].params. */
int lj_autotune_batched(lj_context* ctx, lj_system* sys, int32_t nsys, const float* x,
                        const float* y, const float* z, int32_t reps, lj_tuning* out,
                        void* stream);

#ifdef __cplusplus
}
#endif
#endif
