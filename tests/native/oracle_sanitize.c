/* Sanitizer run of the oracle (oracle/lj_oracle.c) and the engine's host generator
 * (csrc/gen.cpp), built with -fsanitize=address,undefined by tests/native/Makefile and
 * run by tests/test_cpu_sanitize.py.  Every oracle entry point is called on small
 * inputs that reach its edge cases (no particles, one particle, particles outside the
 * box, coincident particles, a subset), and the results are cross-checked so that a
 * run that only "does not crash" still fails on a wrong answer.  Exit status 0 = all
 * checks passed; the sanitizers abort the run on any memory or UB error. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lj_amd.h"
#include "lj_oracle.h"

static int failures = 0;
#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                                  \
        }                                                                \
    } while (0)

static orc_params params(float rc, float lo, float hi, int pred) {
    orc_params p;
    memset(&p, 0, sizeof p);
    p.rc = rc;
    p.sigma = rc / 2.5f;
    p.epsilon = 1.0f;
    p.softening2 = pred == ORC_PRED_CUDALJ ? 1e-5f : 0.0f;
    for (int d = 0; d < 3; ++d) {
        p.box_lo[d] = lo;
        p.box_hi[d] = hi;
    }
    p.predicate = pred;
    return p;
}

static void run_case(const char* name, const float* x, const float* y, const float* z, int64_t n, orc_params p) {
    int32_t dims[3];
    int64_t nc = 0;
    CHECK(orc_grid_dims(&p, dims, &nc) == 0);
    uint32_t* off = malloc(sizeof(uint32_t) * (size_t)(nc + 1));
    uint32_t* perm = malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    uint32_t* keys = malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    CHECK(orc_bin(&p, x, y, z, n, off, perm, keys) == 0);
    CHECK(off[0] == 0 && off[nc] == (uint32_t)n);
    const uint64_t cand = orc_candidates(dims, off);
    CHECK(cand == orc_candidates_direct(&p, x, y, z, n, off));
    uint64_t* rb = malloc(sizeof(uint64_t) * (size_t)(n + 1));
    uint64_t* rg = malloc(sizeof(uint64_t) * (size_t)(n + 1));
    const uint64_t mb = orc_pairs_brute(&p, x, y, z, n, rb, NULL, 0);
    const uint64_t mg = orc_pairs_cellgrid(&p, x, y, z, n, rg, NULL, 0);
    uint32_t* cb = malloc(sizeof(uint32_t) * (size_t)(mb + 1));
    uint32_t* cg = malloc(sizeof(uint32_t) * (size_t)(mg + 1));
    orc_pairs_brute(&p, x, y, z, n, rb, cb, mb);
    orc_pairs_cellgrid(&p, x, y, z, n, rg, cg, mg);
    if (p.predicate != ORC_PRED_CUDALJ) {
        CHECK(mb == mg);
        CHECK(memcmp(rb, rg, sizeof(uint64_t) * (size_t)(n + 1)) == 0);
        CHECK(mb == 0 || memcmp(cb, cg, sizeof(uint32_t) * (size_t)mb) == 0);
    }
    /* every third particle as a subset */
    int64_t ns = (n + 2) / 3;
    int64_t* sub = malloc(sizeof(int64_t) * (size_t)(ns > 0 ? ns : 1));
    for (int64_t r = 0; r < ns; ++r) sub[r] = 3 * r;
    uint64_t* rr = malloc(sizeof(uint64_t) * (size_t)(ns + 1));
    const uint64_t mr = orc_pairs_rows(&p, x, y, z, n, sub, ns, rr, NULL, 0, 2);
    uint32_t* cr = malloc(sizeof(uint32_t) * (size_t)(mr + 1));
    CHECK(orc_pairs_rows(&p, x, y, z, n, sub, ns, rr, cr, mr, 2) == mr);
    for (int64_t r = 0; r < ns; ++r) {
        const uint64_t a = rg[sub[r]], b = rg[sub[r] + 1];
        CHECK(rr[r + 1] - rr[r] == b - a);
        if (rr[r + 1] - rr[r] == b - a && b > a) CHECK(memcmp(cr + rr[r], cg + a, sizeof(uint32_t) * (size_t)(b - a)) == 0);
    }
    double* f64 = malloc(sizeof(double) * 5 * (size_t)(n > 0 ? n : 1));
    double sums[2];
    uint64_t counts[2];
    CHECK(orc_forces_f64(&p, x, y, z, n, NULL, 0, f64, f64 + n, f64 + 2 * n, f64 + 3 * n, f64 + 4 * n, sums, counts, 2) == 0);
    CHECK(counts[0] == cand);
    /* the Triangle model may add a pair more than once (n_pairs counts hits) */
    if (p.predicate == ORC_PRED_TRIANGLE) CHECK(counts[1] >= mg);
    else if (p.predicate != ORC_PRED_CUDALJ) CHECK(counts[1] == mg);
    double* s64 = malloc(sizeof(double) * 5 * (size_t)(ns > 0 ? ns : 1));
    double ssum[2];
    uint64_t scount[2];
    CHECK(orc_forces_f64(&p, x, y, z, n, sub, ns, s64, s64 + ns, s64 + 2 * ns, s64 + 3 * ns, s64 + 4 * ns, ssum, scount, 2) == 0);
    for (int64_t r = 0; r < ns; ++r) CHECK(s64[r] == f64[sub[r]] || (isnan(s64[r]) && isnan(f64[sub[r]])));
    float* f32 = malloc(sizeof(float) * 4 * (size_t)(n > 0 ? n : 1));
    double fsum[2];
    uint64_t fcount[2];
    CHECK(orc_forces_f32_cellgrid(&p, x, y, z, n, f32, f32 + n, f32 + 2 * n, f32 + 3 * n, fsum, fcount, 2) == 0);
    CHECK(fcount[0] == counts[0] && fcount[1] == counts[1]);
    double* ecu = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    CHECK(orc_cudalj_energy(&p, x, y, z, n, ecu) == 0);
    printf("%-12s n=%lld candidates=%llu pairs=%llu\n", name, (long long)n, (unsigned long long)cand,
           (unsigned long long)mg);
    free(off); free(perm); free(keys); free(rb); free(rg); free(cb); free(cg); free(sub); free(rr); free(cr);
    free(f64); free(s64); free(f32); free(ecu);
}

int main(void) {
    /* generator: every kind, empty and small draws, invalid arguments */
    float x[2000], y[2000], z[2000];
    CHECK(lj_generate(LJ_GEN_UNIFORM, 1, 0, 0, NULL, x, y, z) == LJ_OK);
    CHECK(lj_generate(LJ_GEN_UNIFORM, 1, 0, 2000, NULL, x, y, z) == LJ_OK);
    for (int i = 0; i < 2000; ++i) CHECK(x[i] >= 0.0f && x[i] < 1.0f);
    const double sg[5] = {0.5, 0.5, 0.5, 1.0, 0.0};
    CHECK(lj_generate(LJ_GEN_SURFACE, 2, 3, 2000, sg, x, y, z) == LJ_OK);
    const double cl[4] = {7, 0.02, 0.1, 0.9};
    CHECK(lj_generate(LJ_GEN_CLUSTERS, 3, 0, 2000, cl, x, y, z) == LJ_OK);
    CHECK(lj_generate(9, 1, 0, 10, NULL, x, y, z) == LJ_ERR_INVALID_ARG);
    CHECK(lj_generate(LJ_GEN_UNIFORM, 1, 0, -1, NULL, x, y, z) == LJ_ERR_INVALID_ARG);
    CHECK(lj_generate(LJ_GEN_UNIFORM, 1, 0, 5, NULL, NULL, NULL, NULL) == LJ_ERR_INVALID_ARG);
    const double bad[2] = {1.0, 0.0};
    CHECK(lj_generate(LJ_GEN_UNIFORM, 1, 0, 5, bad, x, y, z) == LJ_ERR_INVALID_ARG);

    /* the oracle on every predicate, with clusters, points outside the box, duplicates */
    CHECK(lj_generate(LJ_GEN_CLUSTERS, 4, 0, 1500, cl, x, y, z) == LJ_OK);
    for (int i = 1500; i < 1600; ++i) {  /* outside the box on both sides, on faces */
        x[i] = (i % 2) ? -0.3f : 1.2f;
        y[i] = (float)(i - 1500) * 0.01f;
        z[i] = 0.5f;
    }
    for (int i = 1600; i < 1610; ++i) { x[i] = x[0]; y[i] = y[0]; z[i] = z[0]; }  /* d2 = 0 */
    for (int pred = ORC_PRED_AABB; pred <= ORC_PRED_TRIANGLE; ++pred) {
        char name[32];
        snprintf(name, sizeof name, "pred%d", pred);
        run_case(name, x, y, z, 1610, params(0.05f, 0.0f, 1.0f, pred));
    }
    run_case("empty", x, y, z, 0, params(0.05f, 0.0f, 1.0f, ORC_PRED_AABB));
    run_case("single", x, y, z, 1, params(0.05f, 0.0f, 1.0f, ORC_PRED_AABB));
    run_case("one-cell", x, y, z, 300, params(2.0f, 0.0f, 1.0f, ORC_PRED_AABB));

    /* Philox block function and cell key */
    const uint64_t ctr[4] = {0, 0, 0, 0}, key[2] = {0, 0};
    uint64_t out[4];
    orc_philox4x64_10(ctr, key, out);
    orc_params p = params(0.04f, 0.0f, 1.0f, ORC_PRED_AABB);
    int32_t dims[3];
    int64_t nc;
    CHECK(orc_grid_dims(&p, dims, &nc) == 0 && dims[0] == 25);
    CHECK(orc_cell_key(&p, dims, nextafterf(1.0f, 0.0f), 0.0f, 0.0f) == 24u);
    CHECK(orc_cell_key(&p, dims, NAN, -5.0f, 7.0f) == (uint32_t)((24 * 25 + 0) * 25 + 0));
    p.rc = 0.0f;
    CHECK(orc_grid_dims(&p, dims, &nc) != 0);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("all checks passed\n");
    return 0;
}
