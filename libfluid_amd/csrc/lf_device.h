// lf_device.h -- per-scene kernel constants and device helpers shared by the
// WCSPH kernels.  Included by HIP translation units only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lf {

constexpr int kBlock = 256;                   // threads per workgroup of the fixed-shape kernels
constexpr uint32_t kChunk = 512;              // particle-slot granularity of scene ranges (largest block)
constexpr uint32_t kNoScene = 0xFFFFFFFFu;    // chunk owned by no scene
constexpr uint32_t kMaxParticles = 1u << 28;  // pool slots: 16-byte records addressed with 32-bit offsets
// Stable placement by (cell, id): a bucket of at most kRankDirectMax
// particles is ranked by id directly (each member counts the smaller ids,
// O(size^2) per bucket); a larger one is sorted by id as a segment (LSD
// radix sort, O(size) per pass, k_sort_segments).
constexpr uint32_t kRankDirectMax = 512;
constexpr int kScanPerThread = 4;             // cell-count scan: entries per thread (one 16-byte load)
constexpr int kScanTile = kBlock * kScanPerThread;
constexpr uint32_t kScanDirectTiles = 2048;    // above this many tiles the tile prefixes get their own pass

// Neighbour masks of one radius for 9 stencil rows x nslots particle slots,
// row-major (entry r * nslots + i).  Bit k stands for slot b_r + k of row r.
// A row of at most 32 slots keeps its mask in w32, a row of 33..64 slots in
// w64; longer rows have no mask.
struct RowMasks {
    uint32_t *w32;
    uint64_t *w64;
};

// Per-scene constants: everything a substep needs, derived on the host (in
// double, rounded once to float) from lf_params, plus where the scene lives
// in the pool.  A pool holds one table of these in device memory; each
// workgroup serves one scene and reads its entry with scalar loads.
struct Consts {
    // place in the pool: particle slots [first, first + count) hold the
    // scene's particles (its chunks may hold unused slots after them),
    // buckets [cell_base, cell_base + ncells + 1) its cells and its
    // out-of-grid bucket; `scene` indexes per-scene flags
    uint32_t first, count, cell_base, scene;
    // grid (FluidParamData m_res/m_res2/m_numCells, generalised per axis)
    float inv_cell[3];
    int gx, gy, gz;
    uint32_t ncells;
    double half_box[3];
    // smoothing length
    float h, h2, inv_h, eps;
    // neighbour masks: skin_q = 2.2 * 4096 / h turns a particle's move into
    // the skin fraction (in 1/4096) that covers both particles of a pair
    // moving that far, with 10% to spare (the substep's skin radius follows
    // from the scene's largest move: skin_h2s); whether the normals, forces
    // and XSPH passes walk masks at all (off: every row is scanned)
    float skin_q;
    int use_masks;
    // density + EOS: the candidate loop sums t^3 with t = (h2 - r^2) / h2,
    // so a pair's m poly6(r) is mC6h6 t^3
    float inv_h2;  // 1 / h2
    float mC6h6;   // m * 315/(64 pi h^9) * h2^3
    float rho0, inv_rho0, eos_scale, gamma, gamma_m1, K;
    // gamma and gamma - 1 when they are integers in [1, 15], else 0
    // (tait_pow then multiplies instead of calling powf)
    int gamma_i, gamma_m1_i;
    // normals: lazy_normal = no force term reads this scene's normals (no
    // surface tension), so they are computed only when read
    float hmCg6;  // h * m * (-945/(32 pi h^9))
    int lazy_normal;
    // forces
    float mCgs;  // m * (-45/(pi h^6))
    float nu, sigma, alpha, m2, sigma_m, visc_eps, halfh, coh_off;
    double quarter_inv_psize, alpha_inv_rho0;
    int has_visc, has_tension, has_adhesion;
    // integrator
    double g[3], m, dt, dtdt, vlim, vlim2, alim, alim2;
    double vdamp1, psize, h_d, bd_rho0_d;
    // Near-isolated particles (density below rho0/64: few neighbours, all
    // far out) turn the FP32 cancellation in h^2 - r^2 as r -> h into large
    // relative errors of huge values (1/rho^2).  Pairs touching one take the
    // pair geometry in FP64: hh_d = h^2, mC6_d = m * C6; fragile_S bounds the
    // FP32 sum of t^3 (t = (h2 - r^2) / h2) below which the density is
    // recomputed;
    // fragile_ir is the 1/rho above which a particle is near-isolated,
    // fragile_sr the rho_i + rho_j below which an XSPH pair is.
    double hh_d, mC6_d;
    float fragile_S, fragile_ir, fragile_sr;
    // XSPH
    float xsph_k;  // xsph * 2 m * C6
};

__device__ __forceinline__ float3 f3(float4 v) { return make_float3(v.x, v.y, v.z); }

// ratio^e of the Tait EOS (fluidphysics.cu:78,82).  For an integer exponent
// ei in [1, 15] by binary powering in FP64, rounded once to FP32 (branches on
// ei are uniform: one scene per workgroup); otherwise powf.
__device__ __forceinline__ float tait_pow(float ratio, float e, int ei) {
    if (ei > 0) {
        double x = (double)ratio;
        double p = (ei & 1) ? x : 1.0;
        x *= x;
        if (ei & 2) p *= x;
        if (ei >= 4) {
            x *= x;
            if (ei & 4) p *= x;
            if (ei >= 8) p *= x * x;
        }
        return (float)p;
    }
    return powf(ratio, e);
}

// Squared distance with a fixed FMA order.  The cell-key and pair-predicate
// decisions are made on this exact expression, which the oracle reproduces.
__device__ __forceinline__ float dist2(float4 a, float4 b) {
    float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Table of one lane's flattened neighbour walk: the non-empty pseudo-rows of
// its nine stencil rows, in row order, as (mask word, first slot).  A row of
// at most 32 slots is one pseudo-row; a row of 33..64 slots is two, the low
// and high words of its 64-bit mask, the second starting 32 slots later.
// mw[r] is row r's mask (high word zero for a row of at most 32 slots), b[r]
// its first slot.  put(k, mask, base) stores entry k, for k < kFlatEntries
// only; empty pseudo-rows are skipped without a store, so entries [0, cnt)
// are exactly the non-empty ones.  Returns cnt, the number of non-empty
// pseudo-rows (above kFlatEntries the table is incomplete and must not be
// walked).  Shared by the kernels and a host-side check of the table.
constexpr int kFlatEntries = 16;
template <class Put>
__host__ __device__ __forceinline__ uint32_t flat_table(const uint64_t mw[9], const uint32_t b[9], Put &&put) {
    uint32_t cnt = 0u;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
#pragma unroll
        for (uint32_t half = 0u; half < 2u; ++half) {
            const uint32_t m = (uint32_t)(mw[r] >> (32u * half));
            if (m != 0u) {
                if (cnt < (uint32_t)kFlatEntries) put(cnt, m, b[r] + 32u * half);
                ++cnt;
            }
        }
    }
    return cnt;
}

// Per-scene words of the substep, kSceneWords per scene.
// Word 0 holds one byte per flag, so that the kernels setting and clearing
// different flags store to different bytes (plain stores, no
// read-modify-write) and a reader loads the whole word once: far move: a
// particle moved farther than the XSPH skin masks allow (set by k_forces,
// cleared by k_density); near-isolated: a particle's density is below
// rho0/64 (set by k_density, cleared by k_hash).
constexpr uint32_t kSceneWords = 1u;
constexpr uint32_t kFarMoveByte = 0u, kFragileByte = 1u;
__device__ __forceinline__ void set_flag(uint32_t *flags, uint32_t scene, uint32_t byte, uint8_t v) {
    reinterpret_cast<uint8_t *>(flags + kSceneWords * scene)[byte] = v;
}
__device__ __forceinline__ uint32_t flag_word(const uint32_t *flags, uint32_t scene) {
    return flags[kSceneWords * scene];
}
__device__ __forceinline__ bool flag_set(uint32_t word, uint32_t byte) { return ((word >> (8u * byte)) & 0xFFu) != 0u; }

// XSPH skin.  The substep moves a particle by v dt + a dt^2 / 2 (its
// velocity and acceleration as they enter the substep; the wall projection
// that follows does not lengthen a move that starts inside the walls), so a
// pair within h after the move was within h + 2 max|move| before it.  The
// gather pass reduces each scene's largest move, as the skin fraction
// q / 4096 it needs (rounded up, plus one), into kMoveSlots words of the
// scene's move table, one cache line apart (device-scope atomics on one
// address serialise; each workgroup updates the slot of its index); the
// density and force passes read the slots back (scene_move_q).  The skin
// radius is
// (1 + s) h with s = q / 4096 between kSkinMin and kSkinMax.  A particle
// that still moves farther than skin_move_limit (a wall projection of a
// particle that started outside, or a move above the kSkinMax skin) sets
// the scene's far-move flag and the XSPH pass scans: the result is the same
// either way, only the speed differs.
constexpr uint32_t kSkinQMax = 512u;  // s = 1/8: above kSkinMax
constexpr uint32_t kMoveSlots = 16u, kMoveStride = 32u;  // words per scene: kMoveSlots * kMoveStride
__device__ __forceinline__ uint32_t *move_slot(uint32_t *moves, uint32_t scene, uint32_t k) {
    return moves + ((size_t)scene * kMoveSlots + k) * kMoveStride;
}
constexpr double kSkinMin = 0.01, kSkinMax = 0.1;

// Near-isolated marks: one byte per kNearBlock slots of cell order.  A wave
// of the normal, force or XSPH pass takes the FP64-capable walk only when a
// slot of its lanes' nine rows [b[r], e[r]) lies in a marked block; every
// pair that walk treats specially has a particle the density pass marked.
constexpr uint32_t kNearBlock = 64u;
__device__ __forceinline__ bool rows_near(const uint8_t *__restrict__ near, const uint32_t b[9],
                                          const uint32_t e[9]) {
    // A row's first and last blocks (most rows span one or two) are loaded
    // for all nine rows together; blocks between them only for rows of more
    // than 64 slots.
    uint32_t any = 0u;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        if (b[r] < e[r]) {
            const uint32_t k0 = b[r] / kNearBlock, k1 = (e[r] - 1u) / kNearBlock;
            any |= (uint32_t)near[k0] | (uint32_t)near[k1];
        }
    }
#pragma unroll
    for (int r = 0; r < 9; ++r)
        if (b[r] < e[r])
            for (uint32_t k = b[r] / kNearBlock + 1u; k < (e[r] - 1u) / kNearBlock; ++k) any |= near[k];
    return any != 0u;
}
// A particle's move in the substep, v dt + a dt^2 / 2, is bounded without
// reading a by |v| dt + acc_limit dt^2 / 2: the integrator clamps |a| to
// acc_limit (a larger uploaded acceleration is caught by the far-move check).
__device__ __forceinline__ uint32_t skin_quantum(const Consts &c, float4 v) {
    const float dt = (float)c.dt, ahdt2 = (float)(0.5 * c.dtdt * c.alim);
    const float sv = sqrtf(__builtin_fmaf(v.z, v.z, __builtin_fmaf(v.y, v.y, v.x * v.x)));
    const float q = ceilf(__builtin_fmaf(sv, dt, ahdt2) * c.skin_q) + 1.0f;
    return (uint32_t)fminf(q, (float)kSkinQMax);  // NaN -> kSkinQMax
}
// Skin fraction s (exact in FP32) for the scene's quantised largest move.
__device__ __forceinline__ float skin_frac(uint32_t q) {
    return fminf(fmaxf((float)q * (1.0f / 4096.0f), (float)kSkinMin), (float)kSkinMax);
}
// Squared skin radius of the masks, (1 + s)^2 h2 in FP32: sqrt(h2s) >=
// (1 + s) h (1 - 2e-7) with the roundings of 1 + s, the products and
// h2 = h * h.
__device__ __forceinline__ float skin_h2s(const Consts &c, float s) {
    const float k = 1.0f + s;
    return k * k * c.h2;
}
// The largest move (FP32 positions, FP64 distance) for which masks built
// with skin_h2s(s) still hold every XSPH neighbour: a pair within h after
// the move (FP32 predicate: true distance below h (1 + 1e-6)) was within
// h (1 + 1e-6) + 2 d before it, which the FP32 skin predicate accepts when
// that squared, times (1 + 1e-6), is <= h2s; with sqrt(h2s) taken as
// (1 + s) h (1 - 1e-6) the bound needs no square root.
__device__ __forceinline__ double skin_move_limit(const Consts &c, float s) {
    constexpr double kInvSqrt = 0.99999950000037499969;  // 1 / sqrt(1 + 1e-6)
    const double h = (double)c.h;
    return 0.5 * ((1.0 + (double)s) * h * (1.0 - 1e-6) * kInvSqrt - h * (1.0 + 1e-6)) * (1.0 - 1e-9);
}

// |b - a|^2 in FP64 from FP32 positions: the differences and squares are
// exact, so the result is accurate to FP64 rounding.
__device__ __forceinline__ double dist2_d(float4 a, float4 b) {
    const double dx = (double)b.x - (double)a.x, dy = (double)b.y - (double)a.y, dz = (double)b.z - (double)a.z;
    return __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
}

// Cell coordinate along one axis: floor(x * res) (fluidutil.cu:15-18).
__device__ __forceinline__ float cell_coord(float x, float inv_cell) { return floorf(x * inv_cell); }

// Scene-local bucket of a position: cell key kx + gx*(ky + gy*kz)
// (fluidutil.cu:23-39), or ncells for positions outside the grid (also NaN).
__device__ __forceinline__ uint32_t bucket_of(const Consts &c, float4 p) {
    float fx = cell_coord(p.x, c.inv_cell[0]);
    float fy = cell_coord(p.y, c.inv_cell[1]);
    float fz = cell_coord(p.z, c.inv_cell[2]);
    bool in = fx >= 0.0f && fx < (float)c.gx && fy >= 0.0f && fy < (float)c.gy && fz >= 0.0f && fz < (float)c.gz;
    if (!in) return c.ncells;
    return (uint32_t)fx + (uint32_t)c.gx * ((uint32_t)fy + (uint32_t)c.gy * (uint32_t)fz);
}

__device__ __forceinline__ void cell_xyz(const Consts &c, float4 p, int &cx, int &cy, int &cz) {
    cx = (int)cell_coord(p.x, c.inv_cell[0]);
    cy = (int)cell_coord(p.y, c.inv_cell[1]);
    cz = (int)cell_coord(p.z, c.inv_cell[2]);
}

// Visits every slot j of the clipped 3x3x3 cell stencil around (cx,cy,cz).
// With keys ordered x fastest, the three cells of one x-row are contiguous in
// the sorted arrays, so the stencil is nine contiguous slot ranges.  `start`
// is the scene's own table (pool table + cell_base); a table entry plus
// `shift` is a slot.
template <class F>
__device__ __forceinline__ void for_stencil(const Consts &c, const uint32_t *__restrict__ start, uint32_t shift,
                                            int cx, int cy, int cz, F &&f) {
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < c.gx - 1 ? cx + 1 : c.gx - 1;
    const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy < c.gy - 1 ? cy + 1 : c.gy - 1;
    const int z0 = cz > 0 ? cz - 1 : 0, z1 = cz < c.gz - 1 ? cz + 1 : c.gz - 1;
    for (int z = z0; z <= z1; ++z) {
        for (int y = y0; y <= y1; ++y) {
            const uint32_t row = (uint32_t)(z * c.gy + y) * (uint32_t)c.gx;
            const uint32_t b = start[row + (uint32_t)x0] + shift;
            const uint32_t e = start[row + (uint32_t)x1 + 1u] + shift;
            for (uint32_t j = b; j < e; ++j) f(j);
        }
    }
}

// ---- box boundary [0, box] (fluidutil.cu:85-112) ---------------------------
// Per-particle boundary quantities are evaluated in FP64 on the FP32
// position: the box SDF and its finite-difference normal are ill-conditioned
// near edges and at the adhesion singularity, and FP64 costs nothing at one
// evaluation per particle.  Same formula as the reference: q = |x - b| - b
// with b the box half extent, sdf = min(max(q), 0) + |max(q, 0)|.
__device__ __forceinline__ double box_sdf(const Consts &c, double x, double y, double z) {
    const double qx = fabs(x - c.half_box[0]) - c.half_box[0];
    const double qy = fabs(y - c.half_box[1]) - c.half_box[1];
    const double qz = fabs(z - c.half_box[2]) - c.half_box[2];
    const double inside = fmin(fmax(qx, fmax(qy, qz)), 0.0);
    const double ox = fmax(qx, 0.0), oy = fmax(qy, 0.0), oz = fmax(qz, 0.0);
    return inside + sqrt(ox * ox + oy * oy + oz * oz);
}

// boundaryNormalSDF: normalised central difference of the SDF, step 1e-4f.
__device__ __forceinline__ void box_normal(const Consts &c, double x, double y, double z, double n[3]) {
    const double e = (double)0.0001f;
    const double nx = box_sdf(c, x + e, y, z) - box_sdf(c, x - e, y, z);
    const double ny = box_sdf(c, x, y + e, z) - box_sdf(c, x, y - e, z);
    const double nz = box_sdf(c, x, y, z + e) - box_sdf(c, x, y, z - e);
    const double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
    n[0] = nx * inv;
    n[1] = ny * inv;
    n[2] = nz * inv;
}

// DensityColourOperator (fluidutil.cu:67-80).
__device__ __forceinline__ float3 density_colour(const Consts &c, float rho) {
    float r = rho * c.inv_rho0;
    if (r < 1.0f) return make_float3(0.0f, r, 1.0f - r);
    if (r < 2.0f) return make_float3(r - 1.0f, 2.0f - r, 0.0f);
    return make_float3(1.0f, 0.0f, 0.0f);
}

// Element j of a particle array through a 32-bit byte offset (pool slots are
// capped at kMaxParticles, so 16 j < 2^32): the loads then use a scalar base
// plus a 32-bit vector offset instead of 64-bit address math.
template <class T>
__device__ __forceinline__ T at32(const T *__restrict__ base, uint32_t j) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + j * (uint32_t)sizeof(T));
}

// Raw buffer views (range-checked: a load at or past the end returns zeros).
// Loads through them take a per-lane byte offset and a wave-uniform one
// (soffset), so a candidate loop whose slot offset runs in step across the
// wave advances its address with scalar adds.  0x00020000 is the gfx9
// raw-buffer descriptor word 3 (32-bit data format, no swizzle).
// A plane of n floats.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_buffer(const float *base, uint32_t n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), (short)0, (int)(n * (uint32_t)sizeof(float)),
                                             0x00020000);
}
// Four consecutive floats at byte offset voff + soff (soff wave-uniform;
// dword alignment suffices for buffer loads).
__device__ __forceinline__ float4 load4(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}
// Largest v (v < 2^BITS) over the wave's active lanes, by BITS ballots.
template <int BITS>
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    uint32_t mx = 0u;
#pragma unroll
    for (int b = BITS - 1; b >= 0; --b) {
        const uint32_t t = mx | (1u << b);
        if (__ballot(v >= t)) mx = t;
    }
    return mx;
}

// The scene's quantised largest move of the substep (the XSPH skin): the
// largest of its kMoveSlots slots (uniform addresses: scalar loads).
__device__ __forceinline__ uint32_t scene_move_q(const uint32_t *__restrict__ moves, uint32_t scene) {
    const uint32_t *m = moves + (size_t)scene * kMoveSlots * kMoveStride;
    uint32_t q = 0u;
#pragma unroll
    for (uint32_t k = 0; k < kMoveSlots; ++k) q = q > m[k * kMoveStride] ? q : m[k * kMoveStride];
    return q;
}

// Lane-mask forms of per-lane predicates.  lanes(p) is the wave's mask of
// lanes where p holds (a compare lowers to a single v_cmp into it); combining
// masks is scalar work.  keep takes the mask directly: a per-lane bool
// built from an and of two compares would otherwise be materialised as 0/1
// in a vector register and compared again.
__device__ __forceinline__ uint64_t lanes(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// this lane's bit of mask ? v : 0, one v_cndmask_b32.
__device__ __forceinline__ float keep(float v, uint64_t mask) {
    float r;
    asm("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(r) : "v"(v), "s"(mask));
    return r;
}

// Fast reciprocal / reciprocal square root (v_rcp_f32 / v_rsq_f32, ~1 ulp).
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Workgroup index remapped so that each XCD (which takes workgroups
// round-robin) processes one contiguous range of particles: neighbours of a
// particle then mostly live in the same XCD's L2.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// First slot of this workgroup (BS threads per workgroup over the pool's
// slots) and the constants of the scene that owns it, or null for a chunk no
// scene owns.  Scene ranges are whole chunks and BS divides kChunk, so a
// workgroup never straddles two scenes and the lookup is uniform.
template <int BS>
__device__ __forceinline__ const Consts *block_scene(const Consts *__restrict__ cs,
                                                     const uint32_t *__restrict__ chunk, uint32_t &slot0) {
    static_assert(kChunk % BS == 0, "block size must divide the chunk");
    slot0 = xcd_block() * (uint32_t)BS;
    const uint32_t s = chunk[slot0 / kChunk];
    return s == kNoScene ? nullptr : cs + s;
}

}  // namespace lf
