// Sanitizer run of the engine library's host code (built by tests/native/Makefile with
// AddressSanitizer and UndefinedBehaviorSanitizer on the host side of every source;
// the GPU code is built as usual).  It calls what runs without a GPU: parameter
// validation through lj_grid_dims over valid and invalid parameter sets, the
// generators, the version and error strings, the NULL-argument checks of the device
// entry points and, on a machine with no GPU, the failing context creation.  Exit
// status 0 = all checks passed; the sanitizers abort the run on any error.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "lj_amd.h"

static int failures = 0;
#define CHECK(c)                                                           \
    do {                                                                   \
        if (!(c)) {                                                        \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                                    \
        }                                                                  \
    } while (0)

static lj_params base(float rc) {
    lj_params p;
    std::memset(&p, 0, sizeof p);
    p.rc = rc;
    p.sigma = rc / 2.5f;
    p.epsilon = 1.0f;
    for (int d = 0; d < 3; ++d) {
        p.box_lo[d] = 0.0f;
        p.box_hi[d] = 1.0f;
    }
    return p;
}

int main() {
    int32_t dims[3];
    int64_t nc = 0;
    lj_params p = base(0.04f);
    CHECK(lj_grid_dims(&p, dims, &nc) == LJ_OK && dims[0] == 25 && nc == 15625);
    // every refinement the kernel shape accepts
    for (int r : {0, 1, 2, 4})
        for (int rx : {0, 1, 2, 4, 8, 16})
            for (int b : {0, 64, 128, 256})
                for (int dp : {0, 1, 8, 16, 32}) {
                    lj_params q = base(0.01f);
                    q.refine = (int16_t)r;
                    q.refine_x = (int16_t)rx;
                    q.block = (int16_t)b;
                    q.depth = (int16_t)dp;
                    const int rc = lj_grid_dims(&q, dims, &nc);
                    CHECK(rc == LJ_OK || rc == LJ_ERR_INVALID_ARG);
                }
    // invalid parameter sets: each must be rejected with a status, never a crash
    const float nan = std::numeric_limits<float>::quiet_NaN(), inf = std::numeric_limits<float>::infinity();
    std::vector<lj_params> bad;
    for (float v : {0.0f, -1.0f, nan, inf, 1e-30f}) {
        lj_params q = base(0.04f);
        q.rc = v;
        bad.push_back(q);
    }
    for (float v : {0.0f, nan, -inf}) {
        lj_params q = base(0.04f);
        q.sigma = v;
        bad.push_back(q);
    }
    {
        lj_params q = base(0.04f);
        q.box_hi[1] = -1.0f;
        bad.push_back(q);
        q = base(0.04f);
        q.box_lo[2] = nan;
        bad.push_back(q);
        q = base(0.04f);
        q.predicate = 99;
        bad.push_back(q);
        q = base(0.04f);
        q.precision = -3;
        bad.push_back(q);
        q = base(0.04f);
        q.cell_edge = 0.01f;  // below rc
        bad.push_back(q);
        q = base(0.04f);
        q.softening2 = -1.0f;
        bad.push_back(q);
        q = base(0.04f);
        q.refine = 3;
        bad.push_back(q);
        q = base(0.04f);
        q.block = 100;
        bad.push_back(q);
    }
    for (lj_params& q : bad) CHECK(lj_grid_dims(&q, dims, &nc) != LJ_OK);
    CHECK(lj_grid_dims(nullptr, dims, &nc) == LJ_ERR_INVALID_ARG);
    CHECK(lj_grid_dims(&p, nullptr, &nc) == LJ_ERR_INVALID_ARG);
    // generators
    std::vector<float> x(5000), y(5000), z(5000);
    CHECK(lj_generate(LJ_GEN_UNIFORM, 7, 1, 5000, nullptr, x.data(), y.data(), z.data()) == LJ_OK);
    CHECK(lj_generate(LJ_GEN_SURFACE, 7, 1, 5000, nullptr, x.data(), y.data(), z.data()) == LJ_OK);
    CHECK(lj_generate(LJ_GEN_CLUSTERS, 7, 1, 5000, nullptr, x.data(), y.data(), z.data()) == LJ_OK);
    CHECK(lj_generate(LJ_GEN_UNIFORM, 7, 1, 0, nullptr, nullptr, nullptr, nullptr) == LJ_OK);
    CHECK(lj_generate(42, 7, 1, 10, nullptr, x.data(), y.data(), z.data()) == LJ_ERR_INVALID_ARG);
    CHECK(std::strstr(lj_version(), "gfx950") != nullptr);
    // device entry points reject a NULL context before touching the device
    CHECK(lj_compute(nullptr, x.data(), y.data(), z.data(), 10, &p, x.data(), y.data(), z.data(), nullptr, nullptr,
                     nullptr) == LJ_ERR_INVALID_ARG);
    CHECK(lj_bin(nullptr, x.data(), y.data(), z.data(), 10, &p, nullptr, nullptr, nullptr, nullptr) ==
          LJ_ERR_INVALID_ARG);
    CHECK(lj_lbvh_compute(nullptr, x.data(), y.data(), z.data(), 10, &p, x.data(), y.data(), z.data(), nullptr,
                          nullptr, nullptr) == LJ_ERR_INVALID_ARG);
    CHECK(lj_compute_batched(nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr, nullptr) == LJ_ERR_INVALID_ARG);
    uint64_t gen = 0;
    CHECK(lj_workspace_generation(nullptr, &gen) == LJ_ERR_INVALID_ARG);
    CHECK(lj_context_destroy(nullptr) == LJ_OK);
    CHECK(lj_context_create(nullptr, 0) == LJ_ERR_INVALID_ARG);
    // with no GPU, creation fails cleanly and leaves a message
    lj_context* ctx = nullptr;
    const int rc = lj_context_create(&ctx, 0);
    if (rc == LJ_OK) {
        CHECK(lj_workspace_generation(ctx, &gen) == LJ_OK && gen == 0);
        CHECK(lj_context_destroy(ctx) == LJ_OK);
    } else {
        CHECK(rc == LJ_ERR_NO_DEVICE && ctx == nullptr && std::strlen(lj_last_error(nullptr)) > 0);
    }
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("all checks passed\n");
    return 0;
}
