// lf_scene.hip -- scene object, substep orchestration and the device-facing
// half of the C ABI (include/lf_abi.h).
//
// One scene = one HIP stream + device buffers sized for its particle count
// and grid.  lf_step enqueues work and returns; lf_sync waits and reports
// deferred device errors.  No host synchronisation happens inside a substep
// (the reference blocks on a max-occupancy reduce and four device syncs per
// substep, fluidsystem.cu:319-397; the launch shapes here do not depend on
// cell occupancy).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "lf_device.h"
#include "lf_internal.h"
#include "lf_kernels.h"

using lf::fail;

#define LF_HIP(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(LF_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

constexpr double kPi = 3.141592654;  // CUDART_PI_F as defined by the reference (fluiddefines.h:16)
constexpr double kSkin = 0.1;

struct ProfRec {
    int kind;
    hipEvent_t a, b;
};

}  // namespace

struct lf_scene {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t n = 0, ncells = 0, nbuckets = 0, ntiles = 0;
    lf_params params{};
    int debug = 0;
    bool identity = true;  // state arrays are in particle-id order
    uint64_t steps = 0;
    // state (P0,V0,A0) and the substep's working arrays
    float4 *P0 = nullptr, *P1 = nullptr, *V0 = nullptr, *V1 = nullptr, *V2 = nullptr, *A0 = nullptr, *A1 = nullptr;
    float4 *L0 = nullptr, *L1 = nullptr;  // low-order position words
    float4 *D = nullptr, *N = nullptr, *F = nullptr;
    // packed neighbour data: Q = (x, y, z, 1/density or -1), Pr = pressure,
    // X = (integrated position, density); row masks within h (M) and within
    // the skin radius (MS), 9 rows x n, row-major
    float4 *Q = nullptr, *X = nullptr;
    float *Pr = nullptr;
    lf::RowMasks M{nullptr, nullptr}, MS{nullptr, nullptr};
    uint32_t *keys = nullptr, *ranks = nullptr, *slot_src = nullptr, *slot_id = nullptr;
    uint32_t *counts = nullptr, *start = nullptr, *bsum = nullptr, *err = nullptr;
    uint32_t *far_move = nullptr;  // set when a particle moved farther than the XSPH skin allows
    // profiling
    uint32_t prof_mask = 0;
    std::vector<ProfRec> prof_live;
    std::vector<hipEvent_t> ev_pool;
};

namespace {

template <class T>
int dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(LF_ERR_ALLOC, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    }
    return LF_OK;
}

void free_all(lf_scene *s) {
    void *bufs[] = {s->P0, s->P1, s->V0, s->V1, s->V2, s->A0, s->A1, s->L0, s->L1, s->D, s->N, s->F,
                    s->Q, s->X, s->Pr, s->M.w32, s->M.w64, s->MS.w32, s->MS.w64, s->keys, s->ranks,
                    s->slot_src, s->slot_id, s->counts, s->start, s->bsum, s->err, s->far_move};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto &r : s->prof_live) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : s->ev_pool) (void)hipEventDestroy(e);
    s->prof_live.clear();
    s->ev_pool.clear();
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

// Kernel constants for one substep of length dt (host double -> float once).
lf::Consts make_consts(const lf_params &p, uint32_t ncells, float dt, bool use_masks = true) {
    lf::Consts c;
    std::memset(&c, 0, sizeof(c));
    for (int a = 0; a < 3; ++a) {
        c.inv_cell[a] = p.inv_cell[a];
        c.half_box[a] = 0.5 * (double)p.box[a];
        c.g[a] = (double)p.gravity[a];
    }
    c.gx = (int)p.grid[0];
    c.gy = (int)p.grid[1];
    c.gz = (int)p.grid[2];
    c.ncells = ncells;
    const double h = p.h, m = p.mass;
    const double h6 = std::pow(h, 6.0), h9 = std::pow(h, 9.0);
    const double C6 = 315.0 / (64.0 * kPi * h9);    // poly6 (fluidparams.cu:80)
    const double Cg6 = -945.0 / (32.0 * kPi * h9);  // gradPoly6 (fluidparams.cu:81)
    const double Cgs = -45.0 / (kPi * h6);          // gradSpiky (fluidparams.cu:84)
    c.h = p.h;
    c.h2 = p.h * p.h;  // FP32 product, the pair-predicate threshold (m_h2)
    c.inv_h = (float)(1.0 / h);
    c.eps = FLT_MIN;
    // Skin radius of the XSPH masks: (1 + kSkin) h.  XSPH uses them while
    // the largest move of a substep stays below about kSkin h / 2.
    c.h2s = (float)((1.0 + kSkin) * (1.0 + kSkin) * (double)c.h2);
    // A pair within h after the move (FP32 predicate: true distance below
    // h (1 + 1e-6)) was within h (1 + 1e-6) + 2 d before it, which the FP32
    // skin predicate accepts when that squared, times (1 + 1e-6), is <= h2s.
    c.move_limit = 0.5 * (std::sqrt((double)c.h2s / (1.0 + 1e-6)) - (double)c.h * (1.0 + 1e-6)) * (1.0 - 1e-9);
    c.use_masks = use_masks ? 1 : 0;
    c.mC6 = (float)(m * C6);
    c.rho0 = p.rest_density;
    const double inv_rho0 = p.rest_density == 0.0f ? 1.0 : 1.0 / (double)p.rest_density;
    const double inv_gamma = p.gamma == 0.0f ? 1.0 : 1.0 / (double)p.gamma;
    c.inv_rho0 = (float)inv_rho0;
    c.eos_scale = (float)((double)p.gas_constant * (double)p.rest_density * inv_gamma);
    c.gamma = p.gamma;
    c.gamma_m1 = p.gamma - 1.0f;
    c.K = p.gas_constant;
    c.bd_rho0_d = (double)p.rest_density * (double)0.0039062f;  // fluidkernel.cu:65
    c.h_d = h;
    c.hmCg6 = (float)(h * m * Cg6);
    c.mCgs = (float)(m * Cgs);
    c.nu = p.viscosity;
    c.sigma = p.surface_tension;
    c.alpha = p.adhesion;
    c.m2 = (float)(m * m);
    c.sigma_m = (float)((double)p.surface_tension * m);
    c.visc_eps = (float)(0.01 * (double)c.h2);
    c.halfh = 0.5f * p.h;
    c.coh_off = (float)(h6 * (double)0.0156f);
    c.quarter_inv_psize = 0.25 / (double)p.particle_size;
    c.alpha_inv_rho0 = (double)p.adhesion * inv_rho0;
    c.has_visc = p.viscosity != 0.0f;
    c.has_tension = p.surface_tension != 0.0f;
    c.has_adhesion = p.adhesion != 0.0f;
    c.m = m;
    c.dt = (double)dt;
    c.dtdt = (double)dt * (double)dt;
    c.vlim = (double)p.vel_limit;
    c.vlim2 = c.vlim * c.vlim;
    c.alim = (double)p.acc_limit;
    c.alim2 = c.alim * c.alim;
    c.vdamp1 = 1.0 + (double)p.vdamp;
    c.psize = p.particle_size;
    c.xsph_k = (float)((double)p.xsph_scale * 2.0 * m * C6);
    return c;
}

hipEvent_t take_event(lf_scene *s) {
    if (!s->ev_pool.empty()) {
        hipEvent_t e = s->ev_pool.back();
        s->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Runs `launch(L)`; when kernel class `kind` is profiled, L carries start and
// stop events that the launcher binds to its first and last dispatch.
template <class Fn>
void timed(lf_scene *s, int kind, Fn &&launch) {
    lf::Launch L{s->stream, nullptr, nullptr};
    if (s->prof_mask & (1u << kind)) {
        hipEvent_t a = take_event(s), b = take_event(s);
        if (a && b) {
            L.a = a;
            L.b = b;
            launch(L);
            s->prof_live.push_back(ProfRec{kind, a, b});
            return;
        }
        if (a) s->ev_pool.push_back(a);
        if (b) s->ev_pool.push_back(b);
    }
    launch(L);
}

// Hash, scan, provisional scatter and stable gather of the current state
// (P0,V0,A0) into cell order (P1,V1,A1).
void enqueue_sort(lf_scene *s, const lf::Consts &c) {
    timed(s, LF_K_HASH, [&](const lf::Launch &L) { lf::launch_hash(L, c, s->P0, s->n, s->keys, s->ranks, s->counts); });
    timed(s, LF_K_SCAN, [&](const lf::Launch &L) { lf::launch_scan(L, s->counts, s->nbuckets, s->bsum, s->start); });
    timed(s, LF_K_SCATTER, [&](const lf::Launch &L) {
        lf::launch_scatter(L, s->P0, s->keys, s->ranks, s->start, s->n, s->slot_src, s->slot_id);
    });
    timed(s, LF_K_GATHER, [&](const lf::Launch &L) {
        lf::launch_gather(L, s->slot_src, s->slot_id, s->keys, s->start, s->P0, s->V0, s->A0, s->L0, s->n, s->P1,
                          s->V1, s->A1, s->L1, s->err);
    });
}

// One substep.  Buffer roles: state (P0,V0,A0,L0) -> sorted (P1,V1,A1,L1) ->
// integrated P0, A0, L0 and pre-XSPH velocity V2 -> XSPH velocity V0.
void enqueue_substep(lf_scene *s, const lf::Consts &c) {
    enqueue_sort(s, c);
    timed(s, LF_K_DENSITY, [&](const lf::Launch &L) {
        lf::launch_density(L, c, s->P1, s->start, s->n, s->D, s->Q, s->Pr, s->M, s->MS, s->far_move);
    });
    // Normals feed only the surface-tension term; without it the force walk
    // computes them itself (same bits) and the separate pass is not needed.
    const bool own_normal = !c.has_tension && !(s->debug & LF_DEBUG_NORMALS_PASS);
    if (!own_normal)
        timed(s, LF_K_NORMALS,
              [&](const lf::Launch &L) { lf::launch_normals(L, c, s->Q, s->start, s->M, s->n, s->N); });
    timed(s, LF_K_FORCES, [&](const lf::Launch &L) {
        lf::launch_forces(L, c, own_normal, s->P1, s->Q, s->Pr, s->V1, s->A1, s->L1, s->D, s->N, s->start, s->M,
                          s->n, s->P0, s->V2, s->A0, s->L0, s->X, (s->debug & LF_DEBUG_FORCES) ? s->F : nullptr,
                          s->far_move);
    });
    timed(s, LF_K_XSPH, [&](const lf::Launch &L) {
        lf::launch_xsph(L, c, s->P1, s->X, s->V2, s->start, s->MS, s->far_move, s->n, s->V0);
    });
    s->identity = false;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(LF_ERR_HIP, "%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return LF_OK;
}

// Brings the state back to particle-id order (setup-time only).
int to_id_order(lf_scene *s) {
    if (s->identity) return LF_OK;
    lf::launch_to_id_order(s->stream, s->P0, s->V0, s->A0, s->L0, s->D, s->n, s->P1, s->V1, s->A1, s->L1, s->N);
    if (int rc = check_launch("to_id_order")) return rc;
    std::swap(s->P0, s->P1);
    std::swap(s->V0, s->V1);
    std::swap(s->A0, s->A1);
    std::swap(s->L0, s->L1);
    std::swap(s->D, s->N);
    s->identity = true;
    return LF_OK;
}

}  // namespace

extern "C" {

int lf_device_count(int *count) {
    if (!count) return fail(LF_ERR_INVALID, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = (e == hipSuccess) ? c : 0;
    return LF_OK;
}

int lf_create(const lf_params *p, uint32_t num_particles, lf_scene **out) {
    if (!out) return fail(LF_ERR_INVALID, "lf_create: null output handle");
    *out = nullptr;
    if (int rc = lf_params_check(p)) return rc;
    if (num_particles == 0) return fail(LF_ERR_INVALID, "lf_create: zero particles");
    if (num_particles >= lf::kMaxParticles)
        return fail(LF_ERR_INVALID, "lf_create: %u particles exceed the limit of %u per scene", num_particles,
                    lf::kMaxParticles - 1u);
    lf_scene *s = new (std::nothrow) lf_scene();
    if (!s) return fail(LF_ERR_ALLOC, "lf_create: out of host memory");
    s->params = *p;
    s->n = num_particles;
    s->ncells = p->grid[0] * p->grid[1] * p->grid[2];
    s->nbuckets = s->ncells + 1;
    s->ntiles = lf::scan_tiles(s->nbuckets);
    int rc = LF_OK;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = fail(LF_ERR_HIP, "lf_create: %s", hipGetErrorString(e));
    const size_t n = num_particles;
    if (!rc) rc = dalloc(&s->P0, n);
    if (!rc) rc = dalloc(&s->P1, n);
    if (!rc) rc = dalloc(&s->V0, n);
    if (!rc) rc = dalloc(&s->V1, n);
    if (!rc) rc = dalloc(&s->V2, n);
    if (!rc) rc = dalloc(&s->A0, n);
    if (!rc) rc = dalloc(&s->A1, n);
    if (!rc) rc = dalloc(&s->L0, n);
    if (!rc) rc = dalloc(&s->L1, n);
    if (!rc) rc = dalloc(&s->D, n);
    if (!rc) rc = dalloc(&s->N, n);
    if (!rc) rc = dalloc(&s->Q, n);
    if (!rc) rc = dalloc(&s->X, n);
    if (!rc) rc = dalloc(&s->Pr, n);
    if (!rc) rc = dalloc(&s->M.w32, 9 * n);
    if (!rc) rc = dalloc(&s->M.w64, 9 * n);
    if (!rc) rc = dalloc(&s->MS.w32, 9 * n);
    if (!rc) rc = dalloc(&s->MS.w64, 9 * n);
    if (!rc) rc = dalloc(&s->keys, n);
    if (!rc) rc = dalloc(&s->ranks, n);
    if (!rc) rc = dalloc(&s->slot_src, n);
    if (!rc) rc = dalloc(&s->slot_id, n);
    if (!rc) rc = dalloc(&s->counts, (size_t)s->nbuckets);
    if (!rc) rc = dalloc(&s->start, (size_t)s->nbuckets + 1);
    if (!rc) rc = dalloc(&s->bsum, (size_t)s->ntiles + 1);
    if (!rc) rc = dalloc(&s->err, 1);
    if (!rc) rc = dalloc(&s->far_move, 1);
    if (!rc) {
        // Counts start at zero; afterwards every scan clears them.
        e = hipMemsetAsync(s->counts, 0, sizeof(uint32_t) * s->nbuckets, s->stream);
        if (e == hipSuccess) e = hipMemsetAsync(s->start, 0, sizeof(uint32_t) * (s->nbuckets + 1), s->stream);
        if (e == hipSuccess) e = hipMemsetAsync(s->err, 0, sizeof(uint32_t), s->stream);
        if (e != hipSuccess) rc = fail(LF_ERR_HIP, "lf_create: %s", hipGetErrorString(e));
    }
    if (!rc) {
        lf::launch_init_state(s->stream, s->n, s->P0, s->V0, s->A0, s->L0, s->D);
        rc = check_launch("init_state");
    }
    if (!rc) {
        e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = fail(LF_ERR_HIP, "lf_create: %s", hipGetErrorString(e));
    }
    if (rc) {
        free_all(s);
        delete s;
        return rc;
    }
    *out = s;
    return LF_OK;
}

int lf_destroy(lf_scene *s) {
    if (!s) return LF_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return LF_OK;
}

int lf_num_particles(const lf_scene *s, uint32_t *n) {
    if (!s || !n) return fail(LF_ERR_INVALID, "null argument");
    *n = s->n;
    return LF_OK;
}

int lf_set_params(lf_scene *s, const lf_params *p) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    if (int rc = lf_params_check(p)) return rc;
    for (int a = 0; a < 3; ++a)
        if (p->grid[a] != s->params.grid[a] || p->inv_cell[a] != s->params.inv_cell[a] || p->box[a] != s->params.box[a])
            return fail(LF_ERR_INVALID, "lf_set_params: the grid and box are fixed at lf_create");
    s->params = *p;
    return LF_OK;
}

int lf_get_params(const lf_scene *s, lf_params *p) {
    if (!s || !p) return fail(LF_ERR_INVALID, "null argument");
    *p = s->params;
    return LF_OK;
}

int lf_upload(lf_scene *s, uint32_t first, uint32_t count, const float *pos, const float *vel, const float *acc) {
    if (!s || (!pos && count)) return fail(LF_ERR_INVALID, "lf_upload: null argument");
    if ((uint64_t)first + count > s->n)
        return fail(LF_ERR_INVALID, "lf_upload: ids [%u, %llu) exceed the scene's %u particles", first,
                    (unsigned long long)first + count, s->n);
    if (count == 0) return LF_OK;
    if (int rc = to_id_order(s)) return rc;
    std::vector<float> stage(9 * (size_t)count, 0.0f);
    for (size_t k = 0; k < count; ++k)
        for (int a = 0; a < 3; ++a) {
            stage[9 * k + a] = pos[3 * k + a];
            if (vel) stage[9 * k + 3 + a] = vel[3 * k + a];
            if (acc) stage[9 * k + 6 + a] = acc[3 * k + a];
        }
    float *d_stage = nullptr;
    if (int rc = dalloc(&d_stage, stage.size())) return rc;
    hipError_t e = hipMemcpyAsync(d_stage, stage.data(), stage.size() * sizeof(float), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        lf::launch_upload(s->stream, d_stage, first, count, s->P0, s->V0, s->A0, s->L0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d_stage);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "lf_upload: %s", hipGetErrorString(e));
    return LF_OK;
}

int lf_download(lf_scene *s, int order, float *pos, float *col, float *vel, float *acc, uint32_t *ids) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    if (order != LF_ORDER_CELL && order != LF_ORDER_ID) return fail(LF_ERR_INVALID, "lf_download: bad order %d", order);
    const lf::Consts c = make_consts(s->params, s->ncells, 0.0f);
    float *d_out = nullptr;
    uint32_t *d_ids = nullptr;
    if (int rc = dalloc(&d_out, 12 * (size_t)s->n)) return rc;
    if (int rc = dalloc(&d_ids, (size_t)s->n)) {
        (void)hipFree(d_out);
        return rc;
    }
    lf::launch_export(s->stream, c, s->P0, s->V0, s->A0, s->D, s->n, order == LF_ORDER_ID, d_out, d_ids);
    hipError_t e = hipGetLastError();
    std::vector<float> host(12 * (size_t)s->n);
    std::vector<uint32_t> hid(s->n);
    if (e == hipSuccess) e = hipMemcpyAsync(host.data(), d_out, host.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hid.data(), d_ids, hid.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d_out);
    (void)hipFree(d_ids);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "lf_download: %s", hipGetErrorString(e));
    for (size_t k = 0; k < s->n; ++k) {
        const float *h = &host[12 * k];
        for (int a = 0; a < 3; ++a) {
            if (pos) pos[3 * k + a] = h[a];
            if (col) col[3 * k + a] = h[3 + a];
            if (vel) vel[3 * k + a] = h[6 + a];
            if (acc) acc[3 * k + a] = h[9 + a];
        }
        if (ids) ids[k] = hid[k];
    }
    return LF_OK;
}

int lf_step(lf_scene *s, float dt, uint32_t substeps) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    if (!(std::isfinite(dt) && dt > 0.0f)) return fail(LF_ERR_INVALID, "lf_step: dt must be positive and finite");
    if (substeps == 0) return fail(LF_ERR_INVALID, "lf_step: zero substeps");
    // part_dt = full_dt / substeps (fluidsystem.cu:295); exactly `substeps`
    // substeps run (the reference's float-accumulated loop, :301, can run one
    // fewer).
    const float part = dt / float(substeps);
    const lf::Consts c = make_consts(s->params, s->ncells, part, (s->debug & LF_DEBUG_SCAN) == 0);
    for (uint32_t k = 0; k < substeps; ++k) {
        enqueue_substep(s, c);
        ++s->steps;
    }
    return check_launch("lf_step");
}

int lf_sync(lf_scene *s) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    LF_HIP(hipStreamSynchronize(s->stream));
    uint32_t flag = 0;
    LF_HIP(hipMemcpy(&flag, s->err, sizeof(flag), hipMemcpyDeviceToHost));
    if (flag) {
        LF_HIP(hipMemset(s->err, 0, sizeof(uint32_t)));
        return fail(LF_ERR_OVERFLOW, "a cell (or the out-of-grid bucket) held more than %u particles; "
                                     "its in-cell order was not ranked by id", lf::kRankFixMax);
    }
    return LF_OK;
}

int lf_step_count(const lf_scene *s, uint64_t *count) {
    if (!s || !count) return fail(LF_ERR_INVALID, "null argument");
    *count = s->steps;
    return LF_OK;
}

int lf_set_debug(lf_scene *s, int flags) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    if (flags & ~(LF_DEBUG_FORCES | LF_DEBUG_SCAN | LF_DEBUG_NORMALS_PASS)) return fail(LF_ERR_INVALID, "lf_set_debug: unknown flags %d", flags);
    if ((flags & LF_DEBUG_FORCES) && !s->F)
        if (int rc = dalloc(&s->F, 4 * (size_t)s->n)) return rc;
    s->debug = flags;
    return LF_OK;
}

int lf_get_field(lf_scene *s, int field, void *dst, size_t bytes) {
    if (!s || !dst) return fail(LF_ERR_INVALID, "null argument");
    LF_HIP(hipStreamSynchronize(s->stream));
    const size_t n = s->n;
    const void *src = nullptr;
    size_t need = 0;
    switch (field) {
        case LF_FIELD_KEYS: src = s->keys; need = 4 * n; break;
        case LF_FIELD_START: src = s->start; need = 4 * ((size_t)s->nbuckets + 1); break;
        case LF_FIELD_SORTED_IDS: need = 4 * n; break;
        case LF_FIELD_DENSITY: src = s->D; need = 16 * n; break;
        case LF_FIELD_NORMAL: src = s->N; need = 16 * n; break;
        case LF_FIELD_FORCES:
            if (!s->F) return fail(LF_ERR_STATE, "force components are only kept after lf_set_debug(s, 1)");
            src = s->F;
            need = 64 * n;
            break;
        case LF_FIELD_VEL_HALF: src = s->V2; need = 16 * n; break;
        case LF_FIELD_SORTED_POS: src = s->P1; need = 16 * n; break;
        default: return fail(LF_ERR_INVALID, "unknown field %d", field);
    }
    if (bytes != need) return fail(LF_ERR_INVALID, "field %d needs %zu bytes, got %zu", field, need, bytes);
    if (field == LF_FIELD_SORTED_IDS) {
        std::vector<float4> p(n);
        LF_HIP(hipMemcpy(p.data(), s->P1, 16 * n, hipMemcpyDeviceToHost));
        uint32_t *o = (uint32_t *)dst;
        for (size_t k = 0; k < n; ++k) std::memcpy(&o[k], &p[k].w, 4);
        return LF_OK;
    }
    LF_HIP(hipMemcpy(dst, src, need, hipMemcpyDeviceToHost));
    if (field == LF_FIELD_KEYS) {
        uint32_t *o = (uint32_t *)dst;
        for (size_t k = 0; k < n; ++k)
            if (o[k] == s->ncells) o[k] = LF_NULL_HASH;
    }
    return LF_OK;
}

int lf_neighbours(lf_scene *s, uint32_t *counts, uint32_t *list, uint64_t cap, uint64_t *total) {
    if (!s || !counts) return fail(LF_ERR_INVALID, "null argument");
    const lf::Consts c = make_consts(s->params, s->ncells, 0.0f);
    const uint32_t keep_mask = s->prof_mask;
    s->prof_mask = 0;
    enqueue_sort(s, c);
    s->prof_mask = keep_mask;
    const size_t n = s->n;
    uint32_t *d_cnt = nullptr;
    if (int rc = dalloc(&d_cnt, n)) return rc;
    lf::launch_nb_count(s->stream, c, s->P1, s->start, s->n, d_cnt);
    std::vector<uint32_t> cnt_slot(n);
    std::vector<float4> p1(n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(cnt_slot.data(), d_cnt, 4 * n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(p1.data(), s->P1, 16 * n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d_cnt);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "lf_neighbours: %s", hipGetErrorString(e));
    std::vector<uint32_t> slot_of(n);
    for (size_t k = 0; k < n; ++k) {
        uint32_t id;
        std::memcpy(&id, &p1[k].w, 4);
        slot_of[id] = (uint32_t)k;
    }
    std::vector<uint64_t> off(n + 1, 0);
    for (size_t k = 0; k < n; ++k) off[k + 1] = off[k] + cnt_slot[k];
    const uint64_t pairs = off[n];
    if (total) *total = pairs;
    for (size_t id = 0; id < n; ++id) counts[id] = cnt_slot[slot_of[id]];
    if (!list) return LF_OK;
    if (pairs > cap) return fail(LF_ERR_INVALID, "lf_neighbours: %llu pairs exceed cap %llu",
                                 (unsigned long long)pairs, (unsigned long long)cap);
    uint64_t *d_off = nullptr;
    uint32_t *d_list = nullptr;
    if (int rc = dalloc(&d_off, n + 1)) return rc;
    if (int rc = dalloc(&d_list, (size_t)std::max<uint64_t>(pairs, 1))) {
        (void)hipFree(d_off);
        return rc;
    }
    std::vector<uint32_t> lst(pairs);
    e = hipMemcpyAsync(d_off, off.data(), 8 * (n + 1), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        lf::launch_nb_fill(s->stream, c, s->P1, s->start, s->n, d_off, d_list);
        e = hipGetLastError();
    }
    if (e == hipSuccess && pairs)
        e = hipMemcpyAsync(lst.data(), d_list, 4 * pairs, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d_off);
    (void)hipFree(d_list);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "lf_neighbours: %s", hipGetErrorString(e));
    uint64_t o = 0;
    for (size_t id = 0; id < n; ++id) {
        const size_t k = slot_of[id];
        std::copy(lst.begin() + off[k], lst.begin() + off[k + 1], list + o);
        std::sort(list + o, list + o + cnt_slot[k]);
        o += cnt_slot[k];
    }
    return LF_OK;
}

int lf_profile_enable(lf_scene *s, int mask) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    s->prof_mask = (uint32_t)mask;
    return LF_OK;
}

int lf_profile_read(lf_scene *s, uint64_t *launches, double *ms) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    LF_HIP(hipStreamSynchronize(s->stream));
    uint64_t cnt[LF_K_COUNT] = {0};
    double tot[LF_K_COUNT] = {0};
    for (auto &r : s->prof_live) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess && r.kind >= 0 && r.kind < LF_K_COUNT) {
            cnt[r.kind]++;
            tot[r.kind] += t;
        }
        s->ev_pool.push_back(r.a);
        s->ev_pool.push_back(r.b);
    }
    s->prof_live.clear();
    for (int k = 0; k < LF_K_COUNT; ++k) {
        if (launches) launches[k] = cnt[k];
        if (ms) ms[k] = tot[k];
    }
    return LF_OK;
}

int lf_copy_probe(lf_scene *s, size_t bytes, int repeats, double *ms) {
    if (!s || !ms || bytes == 0 || bytes % 16 || repeats <= 0) return fail(LF_ERR_INVALID, "lf_copy_probe: bad argument");
    float4 *a = nullptr, *b = nullptr;
    if (int rc = dalloc(&a, bytes / 16)) return rc;
    if (int rc = dalloc(&b, bytes / 16)) {
        (void)hipFree(a);
        return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemsetAsync(a, 0, bytes, s->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        const lf::Launch plain{s->stream, nullptr, nullptr};
        lf::launch_copy(plain, a, b, bytes / 16);  // warm-up
        (void)hipEventRecord(e0, s->stream);
        for (int r = 0; r < repeats; ++r) lf::launch_copy(plain, a, b, bytes / 16);
        (void)hipEventRecord(e1, s->stream);
        e = hipStreamSynchronize(s->stream);
    }
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "lf_copy_probe: %s", hipGetErrorString(e));
    *ms = (double)t / repeats;
    return LF_OK;
}

}  // extern "C"
