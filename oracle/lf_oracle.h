/*
 * lf_oracle.h -- FP64 CPU oracle for one WCSPH timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libfluid_amd, include/,
 * the fluidtest driver) links, loads or calls this library.  It is used by
 * tests/ and __graft_entry__.smoke() as the reference the GPU results are
 * compared against, and by bench.py's cpu_baseline leg as the timed CPU
 * restatement.
 *
 * It restates, in double precision and single-threaded C++ (plus an OpenMP
 * build of the same source), the per-substep algorithm of
 * FluidSystem::advance (libfluid/cudasrc/fluidsystem.cu:273-399):
 *
 *   hash -> stable sort by (cell key, particle id) -> cell start table ->
 *   density + Tait EOS -> normals -> forces -> leapfrog integrate ->
 *   XSPH (Jacobi).
 *
 * The integer stages (cell keys, sorted order, cell table, neighbour sets)
 * are evaluated in FP32 exactly like the GPU kernels ("replay" functions), so
 * on the same FP32 state they match bit for bit.  The continuous stages are
 * evaluated in FP64.  See lf_oracle.cpp for the per-function citations.
 */
#ifndef LF_ORACLE_H
#define LF_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scene parameters as the host API stores them (FP32), field for field the
 * same meaning as FluidParamData (libfluid/cudasrc/fluidparams.cuh:16-69),
 * generalised from the unit cube to a box [0,box] with its own grid. */
typedef struct orc_params {
    float rest_density;
    float gas_constant;
    float gamma;
    float mass;
    float particle_size;
    float xsph_scale;
    float surface_tension;
    float viscosity;
    float adhesion;
    float vel_limit;
    float acc_limit;
    float vdamp;
    float gravity[3];
    float h;
    float inv_cell[3];
    uint32_t grid[3];
    float box[3];
} orc_params;

/* Step flags. */
#define ORC_PIN_F32 1u   /* evaluate cell keys and the pair predicate in FP32 on fl32(x) */
#define ORC_XSPH_F32_POS 2u  /* XSPH weights from the integrated positions rounded to FP32, as an
                                    implementation storing positions in FP32 (the reference's float3
                                    arrays) sees them: the FP32-storage floor of the XSPH velocity */

/* Optional per-step outputs.  Every pointer may be NULL.  Per-particle arrays
 * are indexed by particle id; `order` is slot -> id in the new cell order;
 * `start` has ncells+2 entries (bucket ncells holds out-of-grid particles). */
typedef struct orc_out {
    uint32_t *keys;      /* n: cell key of each particle (0xFFFFFFFF = out of grid) */
    uint32_t *order;     /* n */
    uint32_t *start;     /* ncells + 2 */
    double *density;     /* n */
    double *pressure;    /* n */
    double *sound_speed; /* n */
    double *normal;      /* 3n */
    double *f_pressure;  /* 3n */
    double *f_viscosity; /* 3n */
    double *f_tension;   /* 3n */
    double *f_adhesion;  /* 3n */
    double *vel_half;    /* 3n: velocity after the integrator, before XSPH */
    double *sdf_after;   /* n: boundary SDF at the integrated position (before projection) */
} orc_out;

int orc_version(void);
int orc_num_threads(void);
uint32_t orc_num_cells(const orc_params *p);

/* --- FP32 replay stages (bit-exact with the GPU on the same FP32 state) --- */
int orc_hash(const orc_params *p, const float *pos, uint32_t n,
             uint32_t *keys, uint32_t *counts /* ncells+1 or NULL */);
int orc_sort(const uint32_t *keys, uint32_t n, uint32_t *order);
int orc_cell_start(const orc_params *p, const uint32_t *keys, uint32_t n, uint32_t *start);
/* Neighbour sets of the density pass (j != i, r^2 <= h^2 in FP32, j inside the
 * clipped 27-cell stencil of i's cell).  counts[i] = set size; lists are
 * written as a CSR with offsets[i] = sum of counts before i (ids ascending).
 * cap = capacity of `list`.  Returns total pairs or a negative error. */
int64_t orc_neighbours(const orc_params *p, const float *pos, uint32_t n,
                       uint32_t *counts, uint32_t *list, uint64_t cap);
/* Same sets by an O(n^2) scan of all pairs; `stencil` != 0 adds the explicit
 * 27-cell key test, 0 drops it. */
int64_t orc_neighbours_brute(const orc_params *p, const float *pos, uint32_t n, int stencil,
                             uint32_t *counts, uint32_t *list, uint64_t cap);

/* --- FP64 step --- */
/* Advances the state (indexed by particle id) by one substep of length dt.
 * pos/vel/acc are 3n doubles, updated in place. */
int orc_step(const orc_params *p, double dt, uint32_t n,
             double *pos, double *vel, double *acc, uint32_t flags, orc_out *out);

/* --- Individual kernels, for known-answer tests --- */
double orc_poly6(const orc_params *p, double r);
double orc_grad_poly6(const orc_params *p, double r);
double orc_grad_spiky(const orc_params *p, double r);
double orc_cohesion(const orc_params *p, double r);
double orc_boundary_density(const orc_params *p, double d);
double orc_sdf(const orc_params *p, const double *x);
void orc_sdf_normal(const orc_params *p, const double *x, double *n);
void orc_colour(const orc_params *p, double density, double *rgb);
/* orc_colour of n densities into rgb[3n]. */
void orc_colours(const orc_params *p, uint32_t n, const double *density, double *rgb);

#ifdef __cplusplus
}
#endif
#endif
