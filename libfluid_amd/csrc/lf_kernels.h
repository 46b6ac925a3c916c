// lf_kernels.h -- host launchers of the kernels in lf_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lf_device.h"

namespace lf {

// Stream plus optional start/stop events bound to the first/last dispatch of
// a launcher (hipExtLaunchKernelGGL).
struct Launch {
    hipStream_t st;
    hipEvent_t a, b;
};

// Device scene table of a pool: cs[s] are the constants of scene s, chunk[c]
// the scene owning slots [512 c, 512 c + 512) (kNoScene if none); nslots is
// the pool's slot count (a multiple of 512), the stride of its per-slot
// planes.  Stepping launches cover slots [0, nlaunch) (a multiple of 512, at
// most nslots), which hold every active scene.  near holds one byte per
// kNearBlock slots of cell order: cleared by the gather pass, set by the
// density pass where a near-isolated particle sits.
struct PoolView {
    const Consts *cs;
    const uint32_t *chunk;
    uint32_t nslots;
    uint32_t nlaunch;
    uint8_t *near;
};

uint32_t scan_tiles(uint32_t m);
// Block sizes the neighbour passes are built for (128, 256, 512).
bool valid_block_size(int bs);

// Also clears the count of large-bucket segments (nsegs).
void launch_hash(const Launch &L, const PoolView &pv, const float4 *P, uint32_t *keys, uint32_t *ranks,
                 uint32_t *counts, uint32_t *flags, uint32_t *moves, uint32_t *nsegs);
// Exclusive scan of counts[0, m) into start[0, m] (start[m] = total), which
// also clears the counts; bsum holds scan_tiles(m) + 1 words.
void launch_scan(const Launch &L, uint32_t *counts, uint32_t m, uint32_t *bsum, uint32_t *start);
// Provisional slots by atomic rank; then every bucket of more than
// kRankDirectMax particles is sorted by id in place (its segment listed in
// segs, nsegs[0] of them; tmp: one uint2 per pool slot, scratch).
void launch_scatter(const Launch &L, const PoolView &pv, const float4 *P, const uint32_t *keys, const uint32_t *ranks,
                    const uint32_t *start, uint32_t *slot_src, uint32_t *slot_id, uint4 *segs, uint32_t *nsegs,
                    uint2 *tmp);
// Copies positions and velocities to cell order; order[f] receives the state
// slot of the particle in slot f (acceleration and low position words stay
// in the state arrays, read there by the force pass).
void launch_gather(const Launch &L, const PoolView &pv, const uint32_t *slot_src, const uint32_t *slot_id,
                   const uint32_t *keys, const uint32_t *start, const float4 *P, const float4 *V, float4 *Ps,
                   float4 *Vs, uint32_t *order, float *soa, uint32_t *moves);
// Writes one set of neighbour masks, MS: the slots within the substep's
// skin radius (at least h), which the normals and forces passes walk
// (re-testing r <= h) and XSPH walks after the move.
void launch_density(const Launch &L, const PoolView &pv, int bs, const float4 *Ps, const float *soa,
                    const uint32_t *start, float4 *D, float4 *Q, float *Pr, RowMasks MS, uint32_t *flags,
                    const uint32_t *moves);
// The neighbour passes take `flat`: 1 walks the masks as one flattened
// stream of set bits (where the rows fit) rather than row by row (0), 2 does
// so only in waves with a row of more than 32 slots; results are the same
// either way.  launch_normals processes the scenes without lazy_normal,
// or with `all` every scene (the on-demand normals of the last substep).
void launch_normals(const Launch &L, const PoolView &pv, int bs, int flat, bool all, const float4 *Q,
                    const uint32_t *start, RowMasks M, const uint32_t *flags, float4 *N);
// A and Lw are the state arrays, read at order[i]; Aout and Lout must be
// other buffers (they are written in cell order).
void launch_forces(const Launch &L, const PoolView &pv, int bs, int flat, const float4 *Ps, const float4 *Q,
                   const float *Pr, const float4 *Vs, const float4 *A, const float4 *Lw, const uint32_t *order,
                   const float4 *D, const float4 *N, const uint32_t *start, RowMasks M, float4 *Pout, float4 *Vout,
                   float4 *Aout, float4 *Lout, float4 *Fdbg, uint32_t *flags, const uint32_t *moves);
void launch_xsph(const Launch &L, const PoolView &pv, int bs, int flat, const float4 *Ps, const float4 *Pn,
                 const float4 *Vh, const float4 *Lw, const uint32_t *start, RowMasks MS, const uint32_t *flags,
                 float4 *Vout);

// Per-scene helpers: slots [first, first + count) of one scene.
void launch_upload(hipStream_t st, const float *stage, uint32_t first, uint32_t id0, uint32_t count, float4 *P,
                   float4 *V, float4 *A, float4 *Lw);
void launch_init_state(hipStream_t st, uint32_t first, uint32_t count, float4 *P, float4 *V, float4 *A, float4 *Lw,
                       float4 *D);
void launch_to_id_order(hipStream_t st, uint32_t first, uint32_t count, const float4 *P, const float4 *V,
                        const float4 *A, const float4 *Lw, const float4 *D, float4 *P2, float4 *V2, float4 *A2,
                        float4 *L2, float4 *D2);
void launch_move(hipStream_t st, uint32_t from, uint32_t to, uint32_t count, const float4 *P, const float4 *V,
                 const float4 *A, const float4 *Lw, const float4 *D, float4 *P2, float4 *V2, float4 *A2, float4 *L2,
                 float4 *D2);
void launch_carry(hipStream_t st, uint32_t first, uint32_t count, const float4 *A, const float4 *Lw, float4 *A2,
                  float4 *L2);
void launch_export(hipStream_t st, const Consts &c, const float4 *P, const float4 *V, const float4 *A,
                   const float4 *D, bool by_id, float *out, uint32_t *ids);
// Exact state (12 floats per particle: P.xyz, L.xyz, V.xyz, A.xyz), id order.
void launch_state_out(hipStream_t st, const Consts &c, const float4 *P, const float4 *Lw, const float4 *V,
                      const float4 *A, float *out);
void launch_state_in(hipStream_t st, uint32_t first, uint32_t count, const float *in, float4 *P, float4 *Lw,
                     float4 *V, float4 *A);
// Device-side setup of positions (ids id0.., slot first + id).
void launch_setup_axes(hipStream_t st, uint32_t first, uint32_t id0, uint32_t count, uint32_t ny, uint32_t nz,
                       const float *ax, const float *ay, const float *az, float4 *P, float4 *Lw);
void launch_setup_jitter(hipStream_t st, uint32_t first, uint32_t id0, uint32_t nx, uint32_t ny, uint32_t nz,
                         float dx, const float origin[3], float jitter, uint64_t seed, float4 *P, float4 *Lw);
void launch_nb_count(hipStream_t st, const Consts &c, const float4 *Ps, const uint32_t *start, uint32_t *cnt);
void launch_nb_fill(hipStream_t st, const Consts &c, const float4 *Ps, const uint32_t *start, const uint64_t *off,
                    uint32_t *list);
void launch_copy(const Launch &L, const float4 *src, float4 *dst, size_t n4);

}  // namespace lf
