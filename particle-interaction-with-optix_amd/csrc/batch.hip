// batch.hip -- many independent systems in one pipeline of launches.
//
// The reference runs one system per executeSimulation / core call (each with its own
// allocations, launches and host synchronisations, Cuda-version/cudaLJ.cu:515-677).
// Here the cell grids of all systems are concatenated into one key space: system s
// owns global cells [cell_base[s], cell_base[s+1]).  One histogram, one scan, one
// radix sort and one force launch then cover every system, whatever their sizes,
// and one more kernel reduces each system's totals in a fixed order.  Nothing in a
// steady-state call allocates or synchronises with the host, so the whole call can
// be captured in a hipGraph and replayed.
#include <cstring>
#include <vector>

#include "lj_sort.h"
#include "lj_stencil.h"

namespace lj {

namespace {

struct SysDesc {
    GridGeom g;   // force grid (edge cell_edge / refine); cells [cell_base, +g.ncells) of the key space
    GridGeom cg;  // cell_edge grid of the candidate count; cells [ccell_base, +cg.ncells)
    PairConst c;
    Trim tr;
    uint32_t first, n, cell_base, ccell_base;
};

// Largest s in [lo, hi) with base[s] <= v (base ascending, base[lo] <= v).
__device__ inline uint32_t find_sys_in(const uint32_t* __restrict__ base, uint32_t lo, uint32_t hi, uint32_t v) {
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Search hints: hint[j] = the system of value j << kHintShift.  The system of v is
// monotone in v, so it lies between hint[v >> kHintShift] and the next entry: a search
// over one or two systems instead of all of them.  Table of (total >> kHintShift) + 2
// entries for values below `total`.
constexpr int kHintShift = 10;

__host__ __device__ inline size_t hint_entries(uint64_t total) { return (size_t)(total >> kHintShift) + 2; }

__global__ __launch_bounds__(256) void k_build_hints(const uint32_t* __restrict__ base, uint32_t nsys, uint32_t m,
                                                     uint32_t* __restrict__ hint) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint64_t v = (uint64_t)j << kHintShift;
    hint[j] = find_sys_in(base, 0, nsys, v > 0xffffffffull ? 0xffffffffu : (uint32_t)v);
}

// Largest s with base[s] <= v, v below the table's total.
__device__ inline uint32_t find_sys(const uint32_t* __restrict__ base, const uint32_t* __restrict__ hint, uint32_t nsys,
                                    uint32_t v) {
    const uint32_t j = v >> kHintShift;
    const uint32_t lo = hint[j], up = hint[j + 1] + 1;
    return find_sys_in(base, lo, up < nsys ? up : nsys, v);
}

// Global cell key of particle i: its system's first cell plus the cell in that
// system's grid.
struct BatchKeyOp {
    const SysDesc* sys;
    const uint32_t* first;
    const uint32_t* hint;
    uint32_t nsys;
    __device__ uint32_t operator()(float x, float y, float z, uint32_t i) const {
        const SysDesc& d = sys[find_sys(first, hint, nsys, i)];
        return d.cell_base + cell_key(d.g, x, y, z);
    }
};

// Particle counts of the cell_edge grids from the force grids' offsets (refine > 1).
__global__ __launch_bounds__(256) void k_batch_coarse_counts(const uint32_t* __restrict__ foff, uint32_t nccells,
                                                             const SysDesc* __restrict__ sys,
                                                             const uint32_t* __restrict__ ccbase,
                                                             const uint32_t* __restrict__ cchint, uint32_t nsys, int Kx,
                                                             int K, uint32_t* __restrict__ ccount) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nccells) return;
    const SysDesc& d = sys[find_sys(ccbase, cchint, nsys, c)];
    ccount[c] = coarse_count(foff + d.cell_base, d.g, d.cg, Kx, K, c - d.ccell_base);
}

// 27-cell candidates per system on the cell_edge grids (offsets coff over the
// concatenated cell_edge cells), ComputeNbInteractions semantics (cudaLJ.cu:255-297),
// added into cand[system].  A wave's cells usually belong to one system: it then adds
// its sum with one atomic.
__global__ __launch_bounds__(256) void k_batch_candidates(const uint32_t* __restrict__ coff, uint32_t nccells,
                                                          const SysDesc* __restrict__ sys,
                                                          const uint32_t* __restrict__ ccbase,
                                                          const uint32_t* __restrict__ cchint, uint32_t nsys,
                                                          unsigned long long* __restrict__ cand) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t s = 0xffffffffu;
    unsigned long long v = 0;
    if (c < nccells) {
        const uint32_t c0 = coff[c], c1 = coff[c + 1];
        s = find_sys(ccbase, cchint, nsys, c);
        const GridGeom& g = sys[s].cg;
        const uint32_t base = sys[s].ccell_base;
        int32_t cx, cy, cz;
        cell_coords(g, c - base, cx, cy, cz);
        const int32_t gx = g.g[0], gy = g.g[1], gz = g.g[2];
        const int32_t x0 = cx > 0 ? cx - 1 : 0, x1 = cx < gx - 1 ? cx + 1 : gx - 1;
        // The nine rows' ranges loaded at once, with no branch between the loads: a row
        // outside the grid reads the cell's own row and counts zero.  (Loaded row by row
        // under the bounds tests, the loads went out one row at a time.)
        uint32_t lo[9], hi[9];
        bool in[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int32_t zz = cz + q / 3 - 1, yy = cy + q % 3 - 1;
            in[q] = zz >= 0 && zz < gz && yy >= 0 && yy < gy;
            const uint32_t row = base + (uint32_t)(((in[q] ? zz : cz) * gy + (in[q] ? yy : cy)) * gx);
            lo[q] = coff[row + x0];
            hi[q] = coff[row + x1 + 1];
        }
        unsigned long long sum = 0;
#pragma unroll
        for (int q = 0; q < 9; ++q)
            if (in[q]) sum += hi[q] - lo[q];
        const uint32_t nh = c1 - c0;
        if (nh) v = (unsigned long long)nh * (sum - 1);
    }
    const uint32_t s0 = __shfl(s, 0, kWaveSize);
    if (__all(s == s0 || s == 0xffffffffu)) {
        v = wave_sum(v);
        if (__lane_id() == 0 && v) atomicAdd(&cand[s0], v);
    } else if (v) {
        atomicAdd(&cand[s], v);
    }
}

// Energy, virial and accepted pairs of the 64 sorted slots of one wave, when they all
// belong to system `sys`; sys = kSharedWave marks a wave whose slots belong to several
// systems (their values are then kept per slot).
struct WaveTotals {
    double e, v;
    unsigned long long np;
    uint32_t sys, pad;
};
constexpr uint32_t kSharedWave = 0xffffffffu;

// One thread per sorted slot, as the single-system kernel (the same stencil loop and
// LDS queues), with the grid, trimming and pair constants of the slot's system d:
// forces and energies to the caller's order, the slot's energy, virial and accepted
// pairs to ei, vi, np (zero for an idle lane).  Every lane of the wave calls it.
template <int PRED, int PREC, bool QUEUE>
__device__ inline void batch_pairs(const float4* __restrict__ pos, const uint32_t* __restrict__ off, uint32_t k,
                                   bool active, uint32_t key, const SysDesc& d, int Rx, int R, uint32_t depth,
                                   QEntry* __restrict__ queue, float* __restrict__ fx,
                                   float* __restrict__ fy, float* __restrict__ fz, float* __restrict__ e, double& ei,
                                   double& vi, uint32_t& np) {
    const GridGeom g = d.g;
    const PairConst c = d.c;
    const Trim tr = d.tr;
    const uint32_t base = d.cell_base;
    PairSums<PREC, PRED == LJ_PRED_CUDALJ, true> acc;
    const float4 pi = active ? pos[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    acc.set_i(pi);
    acc.allow_exact_diff(c, pi, active);
    np = stencil_pairs<PRED, PREC, true, QUEUE>(pos, off, k, active, pi, g, base, key - base, Rx, R, tr, c, queue,
                                                depth, acc);
    ei = vi = 0.0;
    if (active) {
        const uint32_t i = __float_as_uint(pi.w);
        fx[i] = (float)(-c.e24 * (double)acc.ax.value());
        fy[i] = (float)(-c.e24 * (double)acc.ay.value());
        fz[i] = (float)(-c.e24 * (double)acc.az.value());
        ei = c.e4 * (double)acc.ae.value();
        if (e) e[i] = (float)ei;
        vi = c.e24 * (double)acc.av.value();
    }
}

// The waves whose 64 sorted slots belong to one system (nearly all of them): that
// system's constants are read once per wave (wave-uniform, so they sit in scalar
// registers, not in every lane), and the wave adds its energies, virials and pair
// counts (fixed butterfly order) into one WaveTotals.  A wave shared by several
// systems returns at once; k_batch_force_shared serves it.  A system's sorted slots are
// the caller's index range [first, first + n) of that system (the key space orders
// the systems).
template <int PRED, int PREC, bool QUEUE>
__global__ __launch_bounds__(kMaxBlock) void k_batch_force(const float4* __restrict__ pos,
                                                           const uint32_t* __restrict__ cell_of,
                                                           const uint32_t* __restrict__ off, uint32_t n,
                                                           const SysDesc* __restrict__ sys,
                                                           const uint32_t* __restrict__ cbase,
                                                           const uint32_t* __restrict__ chint, uint32_t nsys, int Rx,
                                                           int R, uint32_t depth, float* __restrict__ fx,
                                                           float* __restrict__ fy, float* __restrict__ fz,
                                                           float* __restrict__ e, WaveTotals* __restrict__ wrec) {
    extern __shared__ QEntry queue[];
    const uint32_t k = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool active = k < n;
    const uint32_t key = active ? cell_of[k] : 0u;
    const uint32_t si = find_sys(cbase, chint, nsys, key);
    // lane 0 holds the wave's lowest slot: idle only when the whole wave is
    const uint32_t s0 = __builtin_amdgcn_readfirstlane(active ? si : kSharedWave);
    if (s0 == kSharedWave || !__all(!active || si == s0)) return;
    double ei, vi;
    uint32_t np;
    batch_pairs<PRED, PREC, QUEUE>(pos, off, k, active, key, sys[s0], Rx, R, depth,
                                   queue + (threadIdx.x / kWaveSize) * depth * kWaveSize + __lane_id(), fx,
                                   fy, fz, e, ei, vi, np);
    const double es = wave_sum(ei), vs = wave_sum(vi);
    const unsigned long long ns = wave_sum((unsigned long long)np);
    if (__lane_id() == 0) wrec[k / kWaveSize] = WaveTotals{es, vs, ns, s0, 0u};
}

// The waves shared by several systems (listed in the plan; at most one per system
// boundary), one per workgroup: each lane reads its own system's constants, and the
// slots' energies, virials and pair counts are kept per slot.
template <int PRED, int PREC, bool QUEUE>
__global__ __launch_bounds__(kWaveSize) void k_batch_force_shared(
    const float4* __restrict__ pos, const uint32_t* __restrict__ cell_of, const uint32_t* __restrict__ off, uint32_t n,
    const SysDesc* __restrict__ sys, const uint32_t* __restrict__ cbase, const uint32_t* __restrict__ chint,
    uint32_t nsys, const uint32_t* __restrict__ shared_waves, int Rx, int R, uint32_t depth, float* __restrict__ fx,
    float* __restrict__ fy, float* __restrict__ fz, float* __restrict__ e, double2* __restrict__ ev,
    uint32_t* __restrict__ npair, WaveTotals* __restrict__ wrec) {
    extern __shared__ QEntry queue[];
    const uint32_t w = shared_waves[blockIdx.x];
    const uint32_t k = w * kWaveSize + threadIdx.x;
    const bool active = k < n;
    const uint32_t key = active ? cell_of[k] : 0u;
    double ei, vi;
    uint32_t np;
    batch_pairs<PRED, PREC, QUEUE>(pos, off, k, active, key, sys[find_sys(cbase, chint, nsys, key)], Rx, R, depth,
                                   queue + (threadIdx.x / kWaveSize) * depth * kWaveSize + __lane_id(), fx, fy, fz, e,
                                   ei, vi, np);
    if (active) {
        ev[k] = make_double2(ei, vi);
        npair[k] = np;
    }
    if (threadIdx.x == 0) wrec[w] = WaveTotals{0.0, 0.0, 0ull, kSharedWave, 0u};
}

// One block per system: FP64 totals over its waves in slot order -- the WaveTotals of
// the waves it fills alone, the per-slot values of its slots in the (at most two)
// waves it shares with other systems -- and its candidate count from cand.
__global__ __launch_bounds__(256) void k_batch_reduce(const SysDesc* __restrict__ sys,
                                                      const WaveTotals* __restrict__ wrec,
                                                      const double2* __restrict__ ev,
                                                      const uint32_t* __restrict__ npair,
                                                      const unsigned long long* __restrict__ cand,
                                                      lj_result* __restrict__ res) {
    const SysDesc& d = sys[blockIdx.x];
    double a = 0.0, b = 0.0;
    unsigned long long np = 0;
    if (d.n) {
        const uint32_t w0 = d.first / kWaveSize, w1 = (d.first + d.n - 1) / kWaveSize;
        for (uint32_t w = w0 + threadIdx.x; w <= w1; w += 256) {
            const WaveTotals r = wrec[w];
            if (r.sys != kSharedWave) {
                a += r.e;
                b += r.v;
                np += r.np;
            } else {
                const uint32_t lo = d.first > w * kWaveSize ? d.first : w * kWaveSize;
                const uint32_t hi = d.first + d.n < (w + 1) * kWaveSize ? d.first + d.n : (w + 1) * kWaveSize;
                for (uint32_t q = lo; q < hi; ++q) {
                    const double2 t = ev[q];
                    a += t.x;
                    b += t.y;
                    np += npair[q];
                }
            }
        }
    }
    block_sum2(a, b);
    const unsigned long long t = block_sum_u64(np);
    if (threadIdx.x == 0) {
        res[blockIdx.x].e_total = 0.5 * a;
        res[blockIdx.x].virial = 0.5 * b;
        res[blockIdx.x].n_pairs = t;
        res[blockIdx.x].n_candidates = cand[blockIdx.x];
    }
}

}  // namespace

}  // namespace lj

using namespace lj;

extern "C" int lj_compute_batched(lj_context* ctx, const lj_system* sys, int32_t nsys, const float* x, const float* y,
                                  const float* z, float* fx, float* fy, float* fz, float* e, lj_result* results,
                                  void* stream) {
    if (!ctx) return LJ_ERR_INVALID_ARG;
    if (nsys < 0 || (nsys > 0 && (!sys || !results))) return fail(ctx, LJ_ERR_INVALID_ARG, "bad system array");
    if (nsys == 0) return LJ_OK;
    // Plan: per-system geometry, contiguous particle ranges, concatenated cells.
    std::vector<SysDesc> desc((size_t)nsys);
    std::vector<uint32_t> first((size_t)nsys), cbase((size_t)nsys), ccbase((size_t)nsys);
    uint64_t ntot = 0, ctot = 0, cctot = 0;
    const lj_params& p0 = sys[0].params;
    for (int32_t s = 0; s < nsys; ++s) {
        const lj_params& ps = sys[s].params;
        GridGeom g;
        PairConst c;
        LJ_TRY(make_geom(ctx, &ps, sys[s].n, &g, &c));
        if (ps.predicate != p0.predicate || ps.precision != p0.precision)
            return fail(ctx, LJ_ERR_INVALID_ARG, "all systems of a batch must share predicate and precision");
        if (ps.refine != p0.refine || ps.refine_x != p0.refine_x || ps.block != p0.block || ps.depth != p0.depth)
            return fail(ctx, LJ_ERR_INVALID_ARG, "all systems of a batch must share refine, refine_x, block and depth");
        if (sys[s].first != (int64_t)ntot)
            return fail(ctx, LJ_ERR_INVALID_ARG, "system %d starts at %lld; systems must be packed back to back from 0",
                        s, (long long)sys[s].first);
        SysDesc& d = desc[(size_t)s];
        std::memset(&d, 0, sizeof d);
        d.g = g;
        d.cg = g;
        d.c = c;
        d.first = (uint32_t)ntot;
        d.n = (uint32_t)sys[s].n;
        d.ccell_base = (uint32_t)cctot;
        ntot += (uint64_t)sys[s].n;
        cctot += g.ncells;
        if (ntot > 0x7fffffffull) return fail(ctx, LJ_ERR_OVERFLOW, "batch holds more than 2^31-1 particles");
        if (cctot > 0x7fffffffull) return fail(ctx, LJ_ERR_OVERFLOW, "batch holds more than 2^31-1 cells");
    }
    // One shape for the batch: the default refinement from the mean occupancy.
    const Shape sh = choose_shape(&p0, (int64_t)ntot, cctot);
    for (int32_t s = 0; s < nsys; ++s) {
        SysDesc& d = desc[(size_t)s];
        if (sh.refine > 1 || sh.refine_x > 1)
            LJ_TRY(refine_geom(ctx, d.cg, &sys[s].params, sh.refine_x, sh.refine, &d.g));
        d.tr = make_trim(&sys[s].params, d.g);
        d.cell_base = (uint32_t)ctot;
        ctot += d.g.ncells;
        if (ctot > 0x7fffffffull) return fail(ctx, LJ_ERR_OVERFLOW, "batch holds more than 2^31-1 force-grid cells");
        first[(size_t)s] = d.first;
        cbase[(size_t)s] = d.cell_base;
        ccbase[(size_t)s] = d.ccell_base;
    }
    if (ntot > 0 && (!x || !y || !z || !fx || !fy || !fz))
        return fail(ctx, LJ_ERR_INVALID_ARG, "positions and forces must not be NULL");
    // waves of 64 sorted slots holding slots of more than one system: the wave of each
    // boundary between non-empty systems that does not fall on a wave boundary
    std::vector<uint32_t> shared;
    {
        bool seen = false;
        for (int32_t s = 0; s < nsys; ++s) {
            if (!desc[(size_t)s].n) continue;
            const uint32_t f = desc[(size_t)s].first;
            if (seen && f % (uint32_t)kWaveSize && (shared.empty() || shared.back() != f / (uint32_t)kWaveSize))
                shared.push_back(f / (uint32_t)kWaveSize);
            seen = true;
        }
    }
    LJ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    BatchWork& W = ctx->batch;
    BinWork& B = ctx->bin;

    // Upload the plan only when it changed (a cached plan keeps replays free of
    // host-to-device copies).
    const size_t dbytes = desc.size() * sizeof(SysDesc), ubytes = (size_t)nsys * 4;
    const size_t sbytes = shared.size() * 4;
    std::vector<uint8_t> blob(dbytes + 3 * ubytes + sbytes);
    std::memcpy(blob.data(), desc.data(), dbytes);
    std::memcpy(blob.data() + dbytes, first.data(), ubytes);
    std::memcpy(blob.data() + dbytes + ubytes, cbase.data(), ubytes);
    std::memcpy(blob.data() + dbytes + 2 * ubytes, ccbase.data(), ubytes);
    if (sbytes) std::memcpy(blob.data() + dbytes + 3 * ubytes, shared.data(), sbytes);
    const SysDesc* dsys = nullptr;
    const uint32_t *dfirst = nullptr, *dcbase = nullptr, *dccbase = nullptr, *dshared = nullptr;
    // search hints of the three base arrays (particles, force-grid cells, cell_edge cells)
    const size_t h1 = hint_entries(ntot), h2 = hint_entries(ctot), h3 = hint_entries(cctot);
    auto bind = [&]() {
        dsys = W.sys.as<SysDesc>();
        const uint8_t* dplan = W.sys.as<uint8_t>();
        dfirst = reinterpret_cast<const uint32_t*>(dplan + dbytes);
        dcbase = reinterpret_cast<const uint32_t*>(dplan + dbytes + ubytes);
        dccbase = reinterpret_cast<const uint32_t*>(dplan + dbytes + 2 * ubytes);
        dshared = reinterpret_cast<const uint32_t*>(dplan + dbytes + 3 * ubytes);
    };
    if (blob != W.plan) {
        LJ_HIP(ctx, W.sys.grow(blob.size()));
        LJ_HIP(ctx, W.hints.grow((h1 + h2 + h3) * sizeof(uint32_t)));
        LJ_HIP(ctx, hipMemcpyAsync(W.sys.ptr, blob.data(), blob.size(), hipMemcpyHostToDevice, s));
        bind();
        uint32_t* hw = W.hints.as<uint32_t>();
        const uint32_t* bases[3] = {dfirst, dcbase, dccbase};
        const size_t sizes[3] = {h1, h2, h3};
        for (int t = 0; t < 3; ++t) {
            hipLaunchKernelGGL(k_build_hints, dim3(div_up((int64_t)sizes[t], 256)), dim3(256), 0, s, bases[t],
                               (uint32_t)nsys, (uint32_t)sizes[t], hw);
            hw += sizes[t];
        }
        LJ_HIP(ctx, hipGetLastError());
        LJ_HIP(ctx, hipStreamSynchronize(s));
        W.plan = std::move(blob);
    }
    bind();
    const uint32_t* hfirst = W.hints.as<uint32_t>();
    const uint32_t* hcbase = hfirst + h1;
    const uint32_t* hccbase = hcbase + h2;

    const int64_t n = (int64_t)ntot;
    const uint32_t ncells = (uint32_t)ctot, nccells = (uint32_t)cctot;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    LJ_HIP(ctx, W.ev.grow(nn * sizeof(double2)));
    LJ_HIP(ctx, W.npair.grow(nn * sizeof(uint32_t)));
    LJ_HIP(ctx, W.wrec.grow(div_up(nn, (size_t)kWaveSize) * sizeof(WaveTotals)));
    uint32_t* off;
    {
        PhaseTimer tb(ctx, LJ_PHASE_BIN, s);
        fill_u64(reinterpret_cast<unsigned long long*>(results), (uint32_t)(4 * nsys), 0ull, s);
        int bits = 0;
        while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)ncells) ++bits;
        LJ_TRY(sort_particles(ctx, x, y, z, n, BatchKeyOp{dsys, dfirst, hfirst, (uint32_t)nsys}, bits, s));
        LJ_TRY(cell_offsets_from_sorted(ctx, n, ncells, s));
        off = B.offsets.as<uint32_t>();
    }
    // The candidate counts need only the cell offsets: they run on the context's second
    // stream beside the force pass, into an array of their own that the final
    // reduction copies into the results after the join.
    LJ_HIP(ctx, W.cand.grow((size_t)nsys * sizeof(unsigned long long)));
    unsigned long long* cand = W.cand.as<unsigned long long>();
    {
        const hipStream_t sc = ctx->aux;
        LJ_HIP(ctx, hipEventRecord(ctx->ev_fork, s));
        LJ_HIP(ctx, hipStreamWaitEvent(sc, ctx->ev_fork, 0));
        fill_u64(cand, (uint32_t)nsys, 0ull, sc);
        const uint32_t* coff = off;
        if (sh.refine > 1 || sh.refine_x > 1) {
            LJ_HIP(ctx, B.ccounts.grow((size_t)nccells * 4));
            LJ_HIP(ctx, B.coffsets.grow(((size_t)nccells + 1) * 4));
            hipLaunchKernelGGL(k_batch_coarse_counts, dim3(div_up(nccells, 256)), dim3(256), 0, sc,
                               (const uint32_t*)off, nccells, dsys, dccbase, hccbase, (uint32_t)nsys, sh.refine_x,
                               sh.refine, B.ccounts.as<uint32_t>());
            uint32_t* co = B.coffsets.as<uint32_t>();
            LJ_TRY(scan_u32(ctx, B.ccounts.as<uint32_t>(), co, nccells, co + nccells, sc));
            coff = co;
        }
        hipLaunchKernelGGL(k_batch_candidates, dim3(div_up(nccells, 256)), dim3(256), 0, sc, coff, nccells, dsys,
                           dccbase, hccbase, (uint32_t)nsys, cand);
        LJ_HIP(ctx, hipEventRecord(ctx->ev_join, sc));
    }
    double2* ev = W.ev.as<double2>();
    uint32_t* npair = W.npair.as<uint32_t>();
    WaveTotals* wrec = W.wrec.as<WaveTotals>();
    const float4* pos = B.pos.as<float4>();
    const uint32_t* cell_of = B.sorted_keys();
    {
        PhaseTimer tf(ctx, LJ_PHASE_FORCE, s);
        const unsigned nb = div_up(n, sh.block);
        const size_t lds = sh.lds_bytes();
        const size_t lds_shared = sh.queued() ? (size_t)sh.depth * kWaveSize * sizeof(QEntry) : 0;
        if (n > 0) {
#define LJ_BFQ(PRED, PREC, Q)                                                                                     \
    do {                                                                                                          \
        hipLaunchKernelGGL((k_batch_force<PRED, PREC, Q>), dim3(nb), dim3(sh.block), lds, s, pos, cell_of,        \
                           (const uint32_t*)off, (uint32_t)n, dsys, dcbase, hcbase, (uint32_t)nsys, sh.refine_x,    \
                           sh.refine, (uint32_t)sh.depth, fx, fy, fz, e, wrec);                                     \
        if (!shared.empty())                                                                                      \
            hipLaunchKernelGGL((k_batch_force_shared<PRED, PREC, Q>), dim3((unsigned)shared.size()),              \
                               dim3(kWaveSize), lds_shared, s, pos, cell_of, (const uint32_t*)off, (uint32_t)n,    \
                               dsys, dcbase, hcbase, (uint32_t)nsys, dshared, sh.refine_x, sh.refine,              \
                               (uint32_t)sh.depth, fx, fy, fz, e, ev, npair, wrec);                                 \
    } while (0)
#define LJ_BF(PRED, PREC)                          \
    do {                                           \
        if (sh.queued()) LJ_BFQ(PRED, PREC, true); \
        else LJ_BFQ(PRED, PREC, false);            \
    } while (0)
#define LJ_BF_PREC(PRED)                                   \
    switch (p0.precision) {                                \
    case LJ_PREC_MIXED: LJ_BF(PRED, LJ_PREC_MIXED); break; \
    case LJ_PREC_FP32: LJ_BF(PRED, LJ_PREC_FP32); break;   \
    default: LJ_BF(PRED, LJ_PREC_FP32_COMP); break;        \
    }
            switch (p0.predicate) {
            case LJ_PRED_AABB: LJ_BF_PREC(LJ_PRED_AABB); break;
            case LJ_PRED_CUSTOM: LJ_BF_PREC(LJ_PRED_CUSTOM); break;
            case LJ_PRED_CUDALJ: LJ_BF_PREC(LJ_PRED_CUDALJ); break;
            case LJ_PRED_SPHERE: LJ_BF_PREC(LJ_PRED_SPHERE); break;
            case LJ_PRED_TRIANGLE: LJ_BF_PREC(LJ_PRED_TRIANGLE); break;
            }
#undef LJ_BF_PREC
#undef LJ_BF
#undef LJ_BFQ
        }
    }
    {
        PhaseTimer to(ctx, LJ_PHASE_OTHER, s);
        LJ_HIP(ctx, hipStreamWaitEvent(s, ctx->ev_join, 0));
        hipLaunchKernelGGL(k_batch_reduce, dim3((unsigned)nsys), dim3(256), 0, s, dsys, (const WaveTotals*)wrec,
                           (const double2*)ev, (const uint32_t*)npair, (const unsigned long long*)cand, results);
    }
    LJ_HIP(ctx, hipGetLastError());
    return LJ_OK;
}
