// lf_noalloc.hip -- lf_step_noalloc: the substeps of one scene on device
// arrays the caller owns, with every working buffer carved from one
// file-scope device array (`scratch`), so that the call allocates nothing:
// no hipMalloc / hipFree, no stream, event or graph creation, no pinned or
// heap host memory.  It is for callers that cannot allocate (work recorded
// under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The work is the same as lf_create + lf_upload + lf_step + lf_download (id
// order) with the default launch shapes, through the same kernels, so the
// results are bitwise those of that sequence.  Each device holds its own
// copy of `scratch` (kScratchBytes) from the library's first call on it
// until the process ends; every call on a device shares it, so calls that
// overlap on one device (other streams, other threads) race.  What must
// start at zero (bucket counts, flags, move slots, marks) is cleared by a
// kernel at the start of every call.

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <utility>

#include "lf_internal.h"
#include "lf_kernels.h"
#include "lf_pool.h"

namespace lf {
namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;  // per device
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Address of `scratch` on each device, taken once per device with that
// device current (each device has its own copy at its own address).  In
// host code the array's name is not a device address; only these are.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

int scratch_base(unsigned char **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "lf_step_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(LF_ERR_INVALID, "lf_step_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(LF_ERR_HIP, "lf_step_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch[dev];
    return LF_OK;
}

// Working buffers of one scene of `slots` particle slots and `buckets`
// buckets (its cells plus the out-of-grid bucket), as the pool lays them
// out, carved in order from a base address with 256-byte alignment.
struct Buffers {
    float4 *P0, *P1, *V0, *V1, *V2, *A0, *A1, *L0, *L1, *D, *N, *Q;
    uint2 *segtmp;
    float *Pr, *soa;
    RowMasks MS;
    uint32_t *keys, *ranks, *slot_src, *slot_id, *start, *bsum, *chunk;
    uint4 *segs;
    Consts *cs;
    // cleared at the start of every call: counts, flags, moves, near, nsegs
    uint32_t *counts, *flags, *moves, *nsegs;
    uint8_t *near;
    uint32_t *zero_begin;
    size_t zero_words;
};

struct Carver {
    unsigned char *base;
    size_t off = 0;
    template <class T>
    T *take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T *p = reinterpret_cast<T *>(base + off);
        off += (count == 0 ? 1 : count) * sizeof(T);
        return p;
    }
};

// Carves the buffers from `base` and returns the bytes used (base may be
// null to size them only).
size_t carve(unsigned char *base, uint32_t slots, uint32_t buckets, Buffers &b) {
    Carver c{base};
    const size_t n = slots;
    // the region cleared at the start of every call, contiguous
    b.counts = c.take<uint32_t>(buckets);
    b.flags = c.take<uint32_t>(kSceneWords);
    b.moves = c.take<uint32_t>((size_t)kMoveSlots * kMoveStride);
    b.nsegs = c.take<uint32_t>(1);
    b.near = c.take<uint8_t>(n / kNearBlock);
    b.zero_begin = b.counts;
    b.zero_words = (c.off - (size_t)((unsigned char *)b.counts - base) + 3) / 4;
    float4 **f4[] = {&b.P0, &b.P1, &b.V0, &b.V1, &b.V2, &b.A0, &b.A1, &b.L0, &b.L1, &b.D, &b.N, &b.Q};
    for (float4 **p : f4) *p = c.take<float4>(n);
    b.segtmp = c.take<uint2>(n);
    b.Pr = c.take<float>(n);
    b.soa = c.take<float>(3 * n);
    b.MS.w32 = c.take<uint32_t>(9 * n + 2);  // +2: an 8-byte read at the last entry stays inside
    b.MS.w64 = c.take<uint64_t>(9 * n);
    uint32_t **u32[] = {&b.keys, &b.ranks, &b.slot_src, &b.slot_id};
    for (uint32_t **p : u32) *p = c.take<uint32_t>(n);
    b.start = c.take<uint32_t>((size_t)buckets + 1);
    b.bsum = c.take<uint32_t>((size_t)scan_tiles(buckets) + 1);
    b.segs = c.take<uint4>(n / (kRankDirectMax + 1) + 1);
    b.cs = c.take<Consts>(1);
    b.chunk = c.take<uint32_t>(n / kChunk);
    return c.off;
}

__global__ __launch_bounds__(kBlock) void k_na_clear(uint32_t *__restrict__ p, size_t words) {
    for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < words; k += (size_t)gridDim.x * kBlock) p[k] = 0u;
}

// The scene table (one entry, passed by value) and its chunk map.
__global__ __launch_bounds__(kBlock) void k_na_tables(Consts c, Consts *__restrict__ cs, uint32_t *__restrict__ chunk,
                                                      uint32_t nchunks) {
    if (threadIdx.x == 0) *cs = c;
    for (uint32_t k = threadIdx.x; k < nchunks; k += kBlock) chunk[k] = 0u;
}

// State of particle id k at slot k, as lf_upload writes it (low position
// words zero).
__global__ __launch_bounds__(kBlock) void k_na_load(const float *__restrict__ pos, const float *__restrict__ vel,
                                                    const float *__restrict__ acc, uint32_t n,
                                                    float4 *__restrict__ P, float4 *__restrict__ V,
                                                    float4 *__restrict__ A, float4 *__restrict__ L) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const size_t o = 3 * (size_t)k;
    P[k] = make_float4(pos[o], pos[o + 1], pos[o + 2], __uint_as_float(k));
    V[k] = make_float4(vel[o], vel[o + 1], vel[o + 2], 0.0f);
    A[k] = make_float4(acc[o], acc[o + 1], acc[o + 2], 0.0f);
    L[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// The state (slot k holds the particle whose id is in P[k].w) back to the
// caller's arrays in id order.
__global__ __launch_bounds__(kBlock) void k_na_store(const float4 *__restrict__ P, const float4 *__restrict__ V,
                                                     const float4 *__restrict__ A, uint32_t n, float *__restrict__ pos,
                                                     float *__restrict__ vel, float *__restrict__ acc) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const float4 p = P[k], v = V[k], a = A[k];
    const size_t o = 3 * (size_t)__float_as_uint(p.w);
    pos[o] = p.x;
    pos[o + 1] = p.y;
    pos[o + 2] = p.z;
    vel[o] = v.x;
    vel[o + 1] = v.y;
    vel[o + 2] = v.z;
    acc[o] = a.x;
    acc[o + 1] = a.y;
    acc[o + 2] = a.z;
}

}  // namespace
}  // namespace lf

extern "C" int lf_step_noalloc(const lf_params *p, uint32_t num_particles, float dt, uint32_t substeps,
                               float *pos_xyz, float *vel_xyz, float *acc_xyz, void *stream) {
    using namespace lf;
    if (int rc = lf_params_check(p)) return rc;
    if (num_particles == 0) return fail(LF_ERR_INVALID, "lf_step_noalloc: zero particles");
    if (!pos_xyz || !vel_xyz || !acc_xyz) return fail(LF_ERR_INVALID, "lf_step_noalloc: null state array");
    if (!(std::isfinite(dt) && dt > 0.0f)) return fail(LF_ERR_INVALID, "lf_step_noalloc: dt must be positive and finite");
    if (substeps == 0) return fail(LF_ERR_INVALID, "lf_step_noalloc: zero substeps");
    const uint64_t cells = (uint64_t)p->grid[0] * p->grid[1] * p->grid[2];
    const uint64_t slots64 = ((uint64_t)num_particles + kChunk - 1) / kChunk * kChunk;
    Buffers b{};
    const size_t need = (slots64 < kMaxParticles && cells + 1 < (1ull << 31))
                            ? carve(nullptr, (uint32_t)slots64, (uint32_t)(cells + 1), b)
                            : ~size_t(0);
    if (need > kScratchBytes)
        return fail(LF_ERR_INVALID,
                    "lf_step_noalloc: %u particles on %llu cells need %zu bytes of scratch; the limit is %zu bytes "
                    "(about 350 bytes per particle, rounded up to 512 particles, plus 8 per cell)",
                    num_particles, (unsigned long long)cells, need, kScratchBytes);
    unsigned char *base = nullptr;
    if (int rc = scratch_base(&base)) return rc;
    const uint32_t slots = (uint32_t)slots64, buckets = (uint32_t)(cells + 1);
    carve(base, slots, buckets, b);
    const hipStream_t st = static_cast<hipStream_t>(stream);

    // part_dt = full_dt / substeps, as lf_step (fluidsystem.cu:295).
    Consts c = make_consts(*p, dt / float(substeps), 0);
    c.first = 0;
    c.count = num_particles;
    c.cell_base = 0;
    c.scene = 0;
    const Tuning t{};
    hipLaunchKernelGGL(k_na_clear, dim3(64), dim3(kBlock), 0, st, b.zero_begin, b.zero_words);
    hipLaunchKernelGGL(k_na_tables, dim3(1), dim3(kBlock), 0, st, c, b.cs, b.chunk, slots / kChunk);
    const unsigned nb = (num_particles + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_na_load, dim3(nb), dim3(kBlock), 0, st, pos_xyz, vel_xyz, acc_xyz, num_particles, b.P0,
                       b.V0, b.A0, b.L0);
    const PoolView pv{b.cs, b.chunk, slots, slots, b.near};
    const Launch L{st, nullptr, nullptr};
    for (uint32_t k = 0; k < substeps; ++k) {
        // One substep, the pool's sequence (Pool::enqueue_sort, enqueue_substep).
        launch_hash(L, pv, b.P0, b.keys, b.ranks, b.counts, b.flags, b.moves, b.nsegs);
        launch_scan(L, b.counts, buckets, b.bsum, b.start);
        launch_scatter(L, pv, b.P0, b.keys, b.ranks, b.start, b.slot_src, b.slot_id, b.segs, b.nsegs, b.segtmp);
        launch_gather(L, pv, b.slot_src, b.slot_id, b.keys, b.start, b.P0, b.V0, b.P1, b.V1, b.ranks, b.soa, b.moves);
        launch_density(L, pv, t.density, b.P1, b.soa, b.start, b.D, b.Q, b.Pr, b.MS, b.flags, b.moves);
        if (!c.lazy_normal)
            launch_normals(L, pv, t.normals, t.flat_normals, false, b.Q, b.start, b.MS, b.flags, b.N);
        launch_forces(L, pv, t.forces, t.flat_forces, b.P1, b.Q, b.Pr, b.V1, b.A0, b.L0, b.ranks, b.D, b.N, b.start,
                      b.MS, b.P0, b.V2, b.A1, b.L1, nullptr, b.flags, b.moves);
        launch_xsph(L, pv, t.xsph, t.flat_xsph, b.P1, b.P0, b.V2, b.L1, b.start, b.MS, b.flags, b.V0);
        std::swap(b.A0, b.A1);
        std::swap(b.L0, b.L1);
    }
    hipLaunchKernelGGL(k_na_store, dim3(nb), dim3(kBlock), 0, st, b.P0, b.V0, b.A0, num_particles, pos_xyz, vel_xyz,
                       acc_xyz);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(LF_ERR_HIP, "lf_step_noalloc: kernel launch failed: %s", hipGetErrorString(e));
    return LF_OK;
}
