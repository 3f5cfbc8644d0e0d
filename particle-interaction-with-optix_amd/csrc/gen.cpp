// gen.cpp -- seeded, counter-based particle generators (host).
//
// Replaces the reference's position generators: cuRAND XORWOW uniforms in
// InitParticles / InitParticlesSphere (Cuda-version/cudaLJ.cu:98-105,174-220) and
// drand48 in core() (SDK/optixMultiSphereAABB/optixSphere.cpp:120-143).  cuRAND is
// not reimplemented; the generator is Philox4x64-10 (Salmon, Moraes, Dror, Shaw,
// SC'11), counter = (particle, stream id, 0, 0), key = (seed, stream).  Every
// particle is an independent function of its index, so host and device, or any
// thread count, produce the same positions.
#include <cmath>
#include <cstdint>

#include "lj_amd.h"

namespace {

struct U4 {
    uint64_t w[4];
};

U4 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0, uint64_t k1) {
    for (int r = 0; r < 10; ++r) {
        const unsigned __int128 p0 = (unsigned __int128)0xD2E7470EE14C6C93ull * c0;
        const unsigned __int128 p1 = (unsigned __int128)0xCA5A826395121157ull * c2;
        const uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
        const uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
        const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B97F4A7C15ull;
        k1 += 0xBB67AE8584CAA73Bull;
    }
    return U4{{c0, c1, c2, c3}};
}

// 24-bit uniform in [0,1): exact in FP32.
inline double u24(uint64_t w) { return (double)(w >> 40) * 0x1p-24; }
// 53-bit uniform in [0,1).
inline double u53(uint64_t w) { return (double)(w >> 11) * 0x1p-53; }
// 53-bit uniform in (0,1] (log-safe).
inline double u53o(uint64_t w) { return (double)((w >> 11) + 1) * 0x1p-53; }

constexpr double kTwoPi = 6.283185307179586476925286766559;

}  // namespace

extern "C" int lj_generate(int kind, uint64_t seed, uint64_t stream, int64_t n, const double* gp, float* x,
                           float* y, float* z) {
    if (n < 0 || (n > 0 && (!x || !y || !z))) return LJ_ERR_INVALID_ARG;
    switch (kind) {
    case LJ_GEN_UNIFORM: {
        const double lo = gp ? gp[0] : 0.0, hi = gp ? gp[1] : 1.0;
        if (!(hi > lo)) return LJ_ERR_INVALID_ARG;
        const float fhi = (float)hi;
        for (int64_t i = 0; i < n; ++i) {
            const U4 r = philox4x64_10((uint64_t)i, 0, 0, 0, seed, stream);
            float v[3];
            for (int d = 0; d < 3; ++d) {
                float f = (float)(lo + u24(r.w[d]) * (hi - lo));
                if (f >= fhi) f = std::nextafter(fhi, (float)lo);
                v[d] = f;
            }
            x[i] = v[0];
            y[i] = v[1];
            z[i] = v[2];
        }
        return LJ_OK;
    }
    case LJ_GEN_SURFACE: {
        // theta = 2 pi u, phi = acos(1 - 2u) (optixSphere.cpp:123-129, cudaLJ.cu:206-215);
        // the optional clamp is cudaLJ's min(2 - 1e-6, .) (cudaLJ.cu:209-215).
        const double cx = gp ? gp[0] : 0.5, cy = gp ? gp[1] : 0.5, cz = gp ? gp[2] : 0.5;
        const double rad = gp ? gp[3] : 1.0, clamp = gp ? gp[4] : 0.0;
        if (!(rad > 0.0)) return LJ_ERR_INVALID_ARG;
        for (int64_t i = 0; i < n; ++i) {
            const U4 r = philox4x64_10((uint64_t)i, 1, 0, 0, seed, stream);
            const double theta = kTwoPi * u53(r.w[0]);
            const double phi = std::acos(1.0 - 2.0 * u53(r.w[1]));
            float v[3] = {(float)(rad * std::sin(phi) * std::cos(theta) + cx),
                          (float)(rad * std::sin(phi) * std::sin(theta) + cy), (float)(rad * std::cos(phi) + cz)};
            if (clamp > 0.0)
                for (float& f : v) f = std::fmin(f, (float)clamp);
            x[i] = v[0];
            y[i] = v[1];
            z[i] = v[2];
        }
        return LJ_OK;
    }
    case LJ_GEN_CLUSTERS: {
        // Isotropic Gaussian clusters, Box-Muller in FP64 (Box & Muller 1958).
        const int64_t ncl = gp ? (int64_t)gp[0] : 256;
        const double sd = gp ? gp[1] : 0.02, cmin = gp ? gp[2] : 0.1, cmax = gp ? gp[3] : 0.9;
        if (ncl < 1 || !(sd >= 0.0) || !(cmax >= cmin)) return LJ_ERR_INVALID_ARG;
        for (int64_t i = 0; i < n; ++i) {
            const int64_t c = i % ncl;
            const U4 rc = philox4x64_10((uint64_t)c, 2, 0, 0, seed, stream);
            const U4 r = philox4x64_10((uint64_t)i, 3, 0, 0, seed, stream);
            const double r1 = std::sqrt(-2.0 * std::log(u53o(r.w[0])));
            const double r2 = std::sqrt(-2.0 * std::log(u53o(r.w[2])));
            const double a1 = kTwoPi * u53(r.w[1]), a2 = kTwoPi * u53(r.w[3]);
            const double g0 = r1 * std::cos(a1), g1 = r1 * std::sin(a1), g2 = r2 * std::cos(a2);
            x[i] = (float)(cmin + (cmax - cmin) * u53(rc.w[0]) + sd * g0);
            y[i] = (float)(cmin + (cmax - cmin) * u53(rc.w[1]) + sd * g1);
            z[i] = (float)(cmin + (cmax - cmin) * u53(rc.w[2]) + sd * g2);
        }
        return LJ_OK;
    }
    default:
        return LJ_ERR_INVALID_ARG;
    }
}
