// ljbench.cpp -- command-line driver with the reference's benchmark sweeps and CSV files.
//
//   ljbench cudalj [-gs|--gensurface] [--max-boxdiv M]
//       Cuda-version/cudaLJ.cu:701-792: for boxDiv = 2,4,..,32 (surface: ..,128; both
//       capped at M) and N = boxDiv^3 * 2^k (k = 0..5), NbLoops = 5 runs of
//       executeSimulation (bin + the 27-cell energy kernel with no cutoff, softening
//       1e-5, sigma = eps = 1); CSV results[-surface]-YYYYmmdd-HHMMSS.csv with the
//       reference columns.
//   ljbench aabb [-rb] [-gs] [-us] [-ns N] [-bd D] [--max-boxdiv M]
//       SDK/optixMultiSphereAABB/optixSphere.cpp:636-830: without -rb one core() call
//       (N = 10, rc = 1/D) whose first 100 energies are checked against a host brute
//       force; with -rb the uniform (boxDiv <= 32) or surface (boxDiv <= 256) sweep,
//       5 core() calls per point, CSV results-spherecustom[sorted][gensurface]-*.csv.
//       The BVH build and traversal are the engine's LBVH (lj_lbvh_compute).
//   ljbench custom|sphere|triangle [-rb] [-gs] [-np|-ns N] [-bd D] [--max-boxdiv M]
//       the Custom, Sphere and Triangle drivers (SDK/optixMultiCustom/optixMultiCustom.cpp:
//       674-861, SDK/optixMultiSphere/optixSphere.cpp:843-1025, SDK/optixMultiTriangle/
//       optixTriangle.cpp:881-1062): the same runs with each program's pair rule (the
//       LJ_PRED_CUSTOM distance test, the SPHERE and TRIANGLE ray-hit models), CSV
//       results-{custom,sphere,triangle}[-surface]-*.csv.
//   ljbench lj --n N --rc RC [--gen uniform|surface|clusters] [--loops L] [--refine K]
//              [--precision mixed|fp32|fp32_comp] [--lbvh] [--seed S]
//       One system with the engine's defaults (sigma = rc/2.5, eps = 1, AABB cutoff,
//       forces + energies); prints pairs/s and the phase times.
//
// Times come from hipEvents on the compute stream (lj_profile_read), so the init
// phase includes the scatter that the reference's timer misses (cudaLJ.cu:635).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <string>
#include <tuple>
#include <vector>

#include "lj_amd.h"

namespace {

#define CHECK_LJ(ctx, expr)                                                                    \
    do {                                                                                      \
        int r__ = (expr);                                                                     \
        if (r__ != LJ_OK) {                                                                   \
            std::fprintf(stderr, "error %d in %s: %s\n", r__, #expr, lj_last_error(ctx));     \
            std::exit(1);                                                                     \
        }                                                                                     \
    } while (0)

#define CHECK_HIP(expr)                                                                       \
    do {                                                                                      \
        hipError_t e__ = (expr);                                                              \
        if (e__ != hipSuccess) {                                                              \
            std::fprintf(stderr, "HIP error %s in %s\n", hipGetErrorString(e__), #expr);      \
            std::exit(1);                                                                     \
        }                                                                                     \
    } while (0)

struct Device {
    float *x = nullptr, *y = nullptr, *z = nullptr, *fx = nullptr, *fy = nullptr, *fz = nullptr, *e = nullptr;
    lj_result* res = nullptr;
    size_t cap = 0;
    void reserve(size_t n) {
        if (n <= cap) return;
        release();
        for (float** p : {&x, &y, &z, &fx, &fy, &fz, &e}) CHECK_HIP(hipMalloc(p, std::max<size_t>(n, 1) * sizeof(float)));
        CHECK_HIP(hipMalloc(&res, sizeof(lj_result)));
        cap = n;
    }
    void release() {
        for (float* p : {x, y, z, fx, fy, fz, e})
            if (p) (void)hipFree(p);
        if (res) (void)hipFree(res);
        x = y = z = fx = fy = fz = e = nullptr;
        res = nullptr;
        cap = 0;
    }
};

lj_params base_params(float rc, float sigma, float eps, float lo, float hi, int pred) {
    lj_params p;
    std::memset(&p, 0, sizeof p);
    p.rc = rc;
    p.sigma = sigma;
    p.epsilon = eps;
    for (int d = 0; d < 3; ++d) {
        p.box_lo[d] = lo;
        p.box_hi[d] = hi;
    }
    p.predicate = pred;
    p.precision = LJ_PREC_MIXED;
    return p;
}

std::string stamp() {
    std::time_t now = std::time(nullptr);
    char buf[80];
    std::strftime(buf, sizeof buf, "%Y%m%d-%H%M%S", std::localtime(&now));
    return buf;
}

struct Timing {
    double init = 0, compute = 0;
};

// One engine call with phase timing; returns the device times in seconds.
Timing timed(lj_context* ctx, bool lbvh, const Device& d, int64_t n, const lj_params& p, hipStream_t s) {
    CHECK_LJ(ctx, lj_profile_enable(ctx, 1));
    if (lbvh)
        CHECK_LJ(ctx, lj_lbvh_compute(ctx, d.x, d.y, d.z, n, &p, d.fx, d.fy, d.fz, d.e, d.res, s));
    else
        CHECK_LJ(ctx, lj_compute(ctx, d.x, d.y, d.z, n, &p, d.fx, d.fy, d.fz, d.e, d.res, s));
    double ms[3];
    unsigned long long calls[3];
    CHECK_LJ(ctx, lj_profile_read(ctx, ms, calls));
    CHECK_LJ(ctx, lj_profile_enable(ctx, 0));
    return {(ms[LJ_PHASE_BIN] + ms[LJ_PHASE_OTHER]) / 1e3, ms[LJ_PHASE_FORCE] / 1e3};
}

void upload(Device& d, const std::vector<float>& x, const std::vector<float>& y, const std::vector<float>& z) {
    const size_t b = x.size() * sizeof(float);
    CHECK_HIP(hipMemcpy(d.x, x.data(), b, hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d.y, y.data(), b, hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d.z, z.data(), b, hipMemcpyHostToDevice));
}

lj_result read_result(const Device& d) {
    lj_result r;
    CHECK_HIP(hipMemcpy(&r, d.res, sizeof r, hipMemcpyDeviceToHost));
    return r;
}

// ------------------------------------------------------------------ cudaLJ mode
struct Frame {
    int nb_particles;
    long long nb_interactions;
    int nb_loops, box_div;
    std::vector<Timing> results;
};

// executeSimulation<float>(N, nbLoops, box, cutoff, checkResult, gensurface)
// (Cuda-version/cudaLJ.cu:515-677): regenerates the particles every loop, bins them,
// counts the 27-cell candidates and runs the energy kernel; returns the summed
// candidate count and the per-loop times.
std::tuple<unsigned long long, std::vector<Timing>> execute_simulation(lj_context* ctx, Device& dev, int n, int loops,
                                                                       float box, float cutoff, bool gensurface) {
    lj_params p = base_params(cutoff, 1.0f, 1.0f, 0.0f, box, LJ_PRED_CUDALJ);
    p.softening2 = 1e-5f;
    int32_t dims[3];
    int64_t nc = 0;
    CHECK_LJ(ctx, lj_grid_dims(&p, dims, &nc));
    std::printf("Start Execution\n - gensurface: %d\n - NbParticles: %d\n - NbLoops: %d\n - Cutoff: %g\n"
                " - Space grid dimension: %d / %d / %d\n - Number of cells: %lld\n - Average particles per cell: %g\n",
                (int)gensurface, n, loops, (double)cutoff, dims[0], dims[1], dims[2], (long long)nc, (double)n / nc);
    dev.reserve((size_t)n);
    std::vector<float> x((size_t)n), y((size_t)n), z((size_t)n);
    std::vector<Timing> out;
    unsigned long long total = 0;
    for (int l = 0; l < loops; ++l) {
        if (gensurface) {
            // radius-1 sphere at (1,1,1), clamped below 2 (InitParticlesSphere, cudaLJ.cu:197-220)
            const double gp[5] = {1.0, 1.0, 1.0, 1.0, 2.0 - 1e-6};
            CHECK_LJ(ctx, lj_generate(LJ_GEN_SURFACE, 1234, (uint64_t)l, n, gp, x.data(), y.data(), z.data()));
        } else {
            const double gp[2] = {0.0, (double)box};
            CHECK_LJ(ctx, lj_generate(LJ_GEN_UNIFORM, 1234, (uint64_t)l, n, gp, x.data(), y.data(), z.data()));
        }
        upload(dev, x, y, z);
        const Timing t = timed(ctx, false, dev, n, p, nullptr);
        const lj_result r = read_result(dev);
        total += r.n_candidates;
        // not printed by the reference: the per-loop counts and the energy total, so a
        // run can be checked against an independent computation on the same positions
        std::printf(" - loop %d: candidates %llu e_total %.17g\n", l, r.n_candidates, r.e_total);
        out.push_back(t);
    }
    double ti = 0, tc = 0;
    for (const Timing& t : out) {
        ti += t.init;
        tc += t.compute;
    }
    std::printf(" - Nb interactions: %llu\n - Time init: %gs\n - Time compute: %gs\n - Total: %gs\n"
                " - GFlops/s: %g\n - Interactions/s: %g\n",
                total, ti, tc, ti + tc, total * 18.0 / (ti + tc) / 1e9, total / (ti + tc));
    return {total, out};
}

int run_cudalj(bool gensurface, int max_boxdiv) {
    lj_context* ctx = nullptr;
    CHECK_LJ(nullptr, lj_context_create(&ctx, 0));
    Device dev;
    std::vector<Frame> all;
    const int loops = 5, max_ppc = 32, max_div = std::min(max_boxdiv, gensurface ? 128 : 32);
    for (int box_div = 2; box_div <= max_div; box_div *= 2) {
        const int nb_boxes = box_div * box_div * box_div;
        for (long long np = nb_boxes; np <= (long long)nb_boxes * max_ppc; np *= 2) {
            const int n = (int)np;
            float box, cell;
            if (gensurface) {
                // cudaLJ.cu:716-746: rc = acos(1 - 2*9*ppc/N), box = ceil(2/rc) rc
                const double ppc = (double)n / nb_boxes;
                const double coef = std::min(1.0, std::max(-1.0, 1.0 - (2 * 9 * ppc) / n));
                const double radius = std::acos(coef);
                box = (float)(std::ceil(2.0 / radius) * radius);
                const int grid = (int)(box / (float)radius);
                cell = box / (float)grid;
            } else {
                box = 1.0f;
                cell = 1.0f / (float)box_div;
            }
            std::printf("NbParticles: %d\nBoxDiv: %d\nCellWidth: %g\nNbLoops: %d\n", n, box_div, (double)cell, loops);
            auto [inter, res] = execute_simulation(ctx, dev, n, loops, box, cell, gensurface);
            all.push_back({n, (long long)inter, loops, box_div, res});
        }
    }
    const std::string fname = std::string("results") + (gensurface ? "-surface" : "") + "-" + stamp() + ".csv";
    std::ofstream f(fname);
    f << "NbParticles,NbInteractions,NbLoops,boxDiv,nbCells,partspercell,interactionsperparticle,timeinit,timecompute,"
         "timetotal\n";
    for (const Frame& fr : all) {
        const long long cells = (long long)fr.box_div * fr.box_div * fr.box_div;
        for (const Timing& t : fr.results)
            f << fr.nb_particles << "," << fr.nb_interactions << "," << fr.nb_loops << "," << fr.box_div << "," << cells
              << "," << (double)fr.nb_particles / cells << "," << (double)fr.nb_interactions / fr.nb_particles << ","
              << t.init << "," << t.compute << "," << t.init + t.compute << "\n";
    }
    std::printf("wrote %s\n", fname.c_str());
    dev.release();
    lj_context_destroy(ctx);
    return 0;
}

// ------------------------------------------------------------------ OptiX program modes
// The four OptiX drivers share one structure and differ in their pair rule, their CSV
// name and the spelling of the particle-count flag.
struct Program {
    const char* name;  // CLI mode
    int predicate;     // the program's pair rule (lj_amd.h)
    bool aabb;         // the AABB driver: -us pre-sort, its own CSV name and checked-run test
};

std::string csv_name(const Program& prog, bool sort, bool gensurface) {
    if (prog.aabb)  // SDK/optixMultiSphereAABB/optixSphere.cpp getFilename
        return std::string("results-spherecustom") + (sort ? "sorted" : "") + (gensurface ? "gensurface" : "") + "-" +
               stamp() + ".csv";
    // SDK/optixMultiCustom/optixMultiCustom.cpp:630-644 and the Sphere / Triangle copies
    return std::string("results-") + prog.name + (gensurface ? "-surface" : "") + "-" + stamp() + ".csv";
}

// core(nbPoints, cutoffRadius, ..., checkResult, [sortParticles,] gensurface)
// (SDK/optixMultiSphereAABB/optixSphere.cpp:82-584, optixMultiCustom.cpp:108-620,
// optixMultiSphere/optixSphere.cpp:94-785, optixMultiTriangle/optixTriangle.cpp:110-825):
// generate, optionally sort into cell order, build the hierarchy, evaluate every
// particle's energy with the program's pair rule (sigma = eps = 1), optionally check
// the first 100 against the driver's host brute force.
std::pair<double, double> core(lj_context* ctx, Device& dev, const Program& prog, int n, float radius, bool check,
                               bool sort, bool gensurface, uint64_t stream) {
    std::vector<float> x((size_t)n), y((size_t)n), z((size_t)n);
    float lo = 0.0f, hi = 1.0f;
    if (gensurface) {
        const double gp[5] = {0.5, 0.5, 0.5, 1.0, 0.0};
        CHECK_LJ(ctx, lj_generate(LJ_GEN_SURFACE, 48, stream, n, gp, x.data(), y.data(), z.data()));
        lo = -0.5f;
        hi = 1.5f;
    } else {
        CHECK_LJ(ctx, lj_generate(LJ_GEN_UNIFORM, 48, stream, n, nullptr, x.data(), y.data(), z.data()));
    }
    lj_params p = base_params(radius, 1.0f, 1.0f, lo, hi, prog.predicate);
    dev.reserve((size_t)n);
    upload(dev, x, y, z);
    double tsort = 0.0;
    if (sort) {
        // the -us pre-sort by cell (sort.cu:28-41), done on the device
        hipEvent_t a, b;
        CHECK_HIP(hipEventCreate(&a));
        CHECK_HIP(hipEventCreate(&b));
        CHECK_HIP(hipEventRecord(a, nullptr));
        CHECK_LJ(ctx, lj_reorder(ctx, dev.x, dev.y, dev.z, n, &p, dev.fx, dev.fy, dev.fz, nullptr, nullptr));
        CHECK_HIP(hipEventRecord(b, nullptr));
        CHECK_HIP(hipEventSynchronize(b));
        float ms = 0;
        CHECK_HIP(hipEventElapsedTime(&ms, a, b));
        tsort = ms / 1e3;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        std::swap(dev.x, dev.fx);
        std::swap(dev.y, dev.fy);
        std::swap(dev.z, dev.fz);
        const size_t bytes = (size_t)n * sizeof(float);
        CHECK_HIP(hipMemcpy(x.data(), dev.x, bytes, hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(y.data(), dev.y, bytes, hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(z.data(), dev.z, bytes, hipMemcpyDeviceToHost));
    }
    const Timing t = timed(ctx, true, dev, n, p, nullptr);
    if (check) {
        std::vector<float> e((size_t)n);
        CHECK_HIP(hipMemcpy(e.data(), dev.e, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
        // checkResult (optixSphere.cpp:550-575; the other drivers test dist < c, e.g.
        // optixMultiCustom.cpp:593-616): host brute force over every other particle
        for (int i = 0; i < std::min(100, n); ++i) {
            double ref = 0.0;
            for (int j = 0; j < n; ++j) {
                if (j == i) continue;
                const double dx = (double)x[j] - x[i], dy = (double)y[j] - y[i], dz = (double)z[j] - z[i];
                const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
                const bool in = prog.aabb ? (d <= radius && d * d >= 1e-16) : d < radius;
                if (in) ref += 4.0 * (std::pow(1.0 / d, 12) - std::pow(1.0 / d, 6));
            }
            std::printf("Energy for particle %d position %.9g %.9g %.9g is %.17g it has been computed as %.9g\n", i,
                        (double)x[i], (double)y[i], (double)z[i], ref, (double)e[(size_t)i]);
        }
    }
    return {tsort + t.init, t.compute};
}

int run_optix(const Program& prog, bool rb, bool gensurface, bool sort, int ns, int bd, int max_div) {
    lj_context* ctx = nullptr;
    CHECK_LJ(nullptr, lj_context_create(&ctx, 0));
    Device dev;
    if (!prog.aabb) sort = false;  // only the AABB driver has the pre-sort
    if (!rb) {
        const float radius = (float)(1.0 / bd);
        std::printf("[LOG] nb spheres = %d\n[LOG] sphere radius = %g\n[LOG] box divisor = %d\n", ns, (double)radius, bd);
        core(ctx, dev, prog, ns, radius, true, prog.aabb, gensurface, 0);
    } else {
        core(ctx, dev, prog, 10, 0.5f, false, sort, gensurface, 0);  // warm-up run, as the reference
        std::vector<Frame> all;
        const int loops = 5, max_ppc = 32;
        const int top = std::min(max_div, gensurface ? 256 : 32);
        for (int box_div = 2; box_div <= top; box_div *= 2) {
            const long long nb_boxes = (long long)box_div * box_div * box_div;
            for (long long np = nb_boxes; np <= nb_boxes * max_ppc; np *= 2) {
                const int n = (int)np;
                float radius;
                if (gensurface) {
                    const double ppc = (double)n / nb_boxes;
                    radius = (float)std::acos(std::min(1.0, std::max(-1.0, 1.0 - (2 * 9 * ppc) / n)));
                } else {
                    radius = 1.0f / (float)box_div;
                }
                std::printf("NbParticles: %d\nBoxDiv: %d\nCellWidth: %g\nNbLoops: %d\n", n, box_div, (double)radius, loops);
                Frame fr{n, 0, loops, box_div, {}};
                // the reference's formula N * (N / nbBoxes) * 27 (optixSphere.cpp:741), in 64 bits
                fr.nb_interactions = (long long)n * (n / nb_boxes) * 27;
                for (int l = 0; l < loops; ++l) {
                    // (the Custom and Triangle surface sweeps of the reference pass the -np
                    // count instead of N here, optixMultiCustom.cpp:769; every point runs N)
                    auto [ti, tc] = core(ctx, dev, prog, n, radius, false, sort, gensurface, (uint64_t)(l + 1));
                    fr.results.push_back({ti, tc});
                }
                all.push_back(fr);
            }
        }
        const std::string fname = csv_name(prog, sort, gensurface);
        std::ofstream f(fname);
        f << "NbParticles,NbInteractions,NbLoops,boxDiv,nbCells,partspercell,timeinit,timecompute,timetotal\n";
        for (const Frame& fr : all) {
            const long long cells = (long long)fr.box_div * fr.box_div * fr.box_div;
            for (const Timing& t : fr.results)
                f << fr.nb_particles << "," << fr.nb_interactions << "," << fr.nb_loops << "," << fr.box_div << ","
                  << cells << "," << (double)fr.nb_particles / cells << "," << t.init << "," << t.compute << ","
                  << t.init + t.compute << "\n";
        }
        std::printf("wrote %s\n", fname.c_str());
    }
    dev.release();
    lj_context_destroy(ctx);
    return 0;
}

// ------------------------------------------------------------------ engine mode
int run_lj(int argc, char** argv) {
    long long n = 100000;
    double rc = 0.04;
    std::string gen = "uniform", prec = "mixed";
    int loops = 5, refine = 0, refine_x = 0, block = 0, depth = 0;
    bool lbvh = false, tune = false;
    uint64_t seed = 1234;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "%s needs a value\n", a.c_str());
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "--n") n = std::atoll(next());
        else if (a == "--rc") rc = std::atof(next());
        else if (a == "--gen") gen = next();
        else if (a == "--loops") loops = std::atoi(next());
        else if (a == "--refine") refine = std::atoi(next());
        else if (a == "--refine-x") refine_x = std::atoi(next());
        else if (a == "--block") block = std::atoi(next());
        else if (a == "--depth") depth = std::atoi(next());
        else if (a == "--autotune") tune = true;
        else if (a == "--precision") prec = next();
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--lbvh") lbvh = true;
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 1;
        }
    }
    lj_context* ctx = nullptr;
    CHECK_LJ(nullptr, lj_context_create(&ctx, 0));
    std::vector<float> x((size_t)n), y((size_t)n), z((size_t)n);
    float lo = 0.0f, hi = 1.0f;
    if (gen == "surface") {
        CHECK_LJ(ctx, lj_generate(LJ_GEN_SURFACE, seed, 0, n, nullptr, x.data(), y.data(), z.data()));
        lo = -0.5f;
        hi = 1.5f;
    } else if (gen == "clusters") {
        CHECK_LJ(ctx, lj_generate(LJ_GEN_CLUSTERS, seed, 0, n, nullptr, x.data(), y.data(), z.data()));
    } else {
        CHECK_LJ(ctx, lj_generate(LJ_GEN_UNIFORM, seed, 0, n, nullptr, x.data(), y.data(), z.data()));
    }
    lj_params p = base_params((float)rc, (float)(rc / 2.5), 1.0f, lo, hi, LJ_PRED_AABB);
    p.refine = (int16_t)refine;
    p.refine_x = (int16_t)refine_x;
    p.block = (int16_t)block;
    p.depth = (int16_t)depth;
    p.precision = prec == "fp32" ? LJ_PREC_FP32 : prec == "fp32_comp" ? LJ_PREC_FP32_COMP : LJ_PREC_MIXED;
    Device dev;
    dev.reserve((size_t)n);
    upload(dev, x, y, z);
    if (tune && !lbvh) {
        lj_tuning t;
        CHECK_LJ(ctx, lj_autotune(ctx, dev.x, dev.y, dev.z, n, &p, 3, &t, nullptr));
        std::printf("autotune: refine %d, refine_x %d, block %d, depth %d (%d shapes; %.4f ms vs default %.4f ms)\n",
                    t.refine, t.refine_x, t.block, t.depth, t.n_tried, t.ms_best, t.ms_default);
    }
    (void)timed(ctx, lbvh, dev, n, p, nullptr);  // warm-up (grows the workspace)
    double ti = 0, tc = 0;
    for (int l = 0; l < loops; ++l) {
        const Timing t = timed(ctx, lbvh, dev, n, p, nullptr);
        ti += t.init;
        tc += t.compute;
    }
    const lj_result r = read_result(dev);
    std::printf("n=%lld rc=%g %s%s: pairs=%llu candidates=%llu e_total=%.17g virial=%.17g\n", n, rc, gen.c_str(),
                lbvh ? " (lbvh)" : "", r.n_pairs, r.n_candidates, r.e_total, r.virial);
    std::printf("per loop: init %.4f ms, force %.4f ms, %.4g pair interactions/s (force kernel %.4g)\n",
                ti / loops * 1e3, tc / loops * 1e3, r.n_pairs / ((ti + tc) / loops), r.n_pairs / (tc / loops));
    dev.release();
    lj_context_destroy(ctx);
    return 0;
}

void usage(const char* argv0) {
    std::fprintf(stderr,
                 "Usage: %s cudalj [-gs|--gensurface] [--max-boxdiv M]\n"
                 "       %s aabb [-rb|--runbench] [-gs|--gensurface] [-us|--usedsort] [-ns|--nbspheres N]\n"
                 "            [-bd|--boxdivisor D] [--max-boxdiv M]\n"
                 "       %s custom|triangle [-rb] [-gs] [-np|--nbpoints N] [-bd D] [--max-boxdiv M]\n"
                 "       %s sphere [-rb] [-gs] [-ns|--nbspheres N] [-bd D] [--max-boxdiv M]\n"
                 "       %s lj [--n N] [--rc RC] [--gen uniform|surface|clusters] [--loops L] [--refine K]\n"
                 "            [--refine-x KX]\n"
                 "            [--block B] [--depth D] [--autotune] [--precision mixed|fp32|fp32_comp] [--lbvh]\n"
                 "            [--seed S]\n",
                 argv0, argv0, argv0, argv0, argv0);
    std::exit(1);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) usage(argv[0]);
    const std::string mode = argv[1];
    if (mode == "lj") return run_lj(argc, argv);
    bool gensurface = false, rb = false, sort = false;
    int ns = 10, bd = 2, max_div = 256;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-gs" || a == "--gensurface") gensurface = true;
        else if (a == "-rb" || a == "--runbench") rb = true;
        else if (a == "-us" || a == "--usedsort") sort = true;
        else if ((a == "-ns" || a == "--nbspheres" || a == "-np" || a == "--nbpoints") && i + 1 < argc)
            ns = std::atoi(argv[++i]);
        else if ((a == "-bd" || a == "--boxdivisor") && i + 1 < argc) bd = std::atoi(argv[++i]);
        else if (a == "--max-boxdiv" && i + 1 < argc) max_div = std::atoi(argv[++i]);
        else if (a == "-h" || a == "--help") usage(argv[0]);
        else {
            std::fprintf(stderr, "Unknown option '%s'\n", a.c_str());
            usage(argv[0]);
        }
    }
    if (mode == "cudalj") return run_cudalj(gensurface, max_div);
    static const Program programs[] = {{"aabb", LJ_PRED_AABB, true},
                                        {"custom", LJ_PRED_CUSTOM, false},
                                        {"sphere", LJ_PRED_SPHERE, false},
                                        {"triangle", LJ_PRED_TRIANGLE, false}};
    for (const Program& prog : programs)
        if (mode == prog.name) return run_optix(prog, rb, gensurface, sort, ns, bd, max_div);
    usage(argv[0]);
}
