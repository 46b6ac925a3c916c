// lf_pool.hip -- pool of scenes: device arrays, scene admission, substep
// orchestration (eager or hipGraph), per-scene data movement, inspection and
// the block-size autotuner.
//
// No host synchronisation happens inside a substep (the reference blocks on
// a max-occupancy reduce and four device syncs per substep,
// fluidsystem.cu:319-397); launch shapes depend only on the pool's capacity.

#include "lf_pool.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <atomic>
#include <numeric>

#include "lf_internal.h"

#define LF_HIP(expr)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return lf::fail(LF_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace lf {

namespace {

std::atomic<int64_t> g_live_allocs{0};  // device allocations the library holds
std::atomic<int64_t> g_fail_at{-1};     // allocations to go before one fails (-1: never)

}  // namespace

int dev_alloc(void **p, size_t bytes) {
    *p = nullptr;
    int64_t k = g_fail_at.load();
    while (k >= 0 && !g_fail_at.compare_exchange_weak(k, k - 1)) {
    }
    if (k == 0) return fail(LF_ERR_ALLOC, "hipMalloc of %zu bytes failed: injected failure", bytes);
    hipError_t e = hipMalloc(p, bytes == 0 ? 1 : bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(LF_ERR_ALLOC, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    g_live_allocs.fetch_add(1);
    return LF_OK;
}

void debug_fail_alloc(int64_t nth) { g_fail_at.store(nth < 0 ? -1 : nth); }
int64_t debug_live_allocs() { return g_live_allocs.load(); }

void dev_free_raw(void *p) {
    if (!p) return;
    (void)hipFree(p);
    g_live_allocs.fetch_sub(1);
}

namespace {

constexpr double kPi = 3.141592654;  // CUDART_PI_F as defined by the reference (fluiddefines.h:16)

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(LF_ERR_HIP, "%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return LF_OK;
}

}  // namespace

// Kernel constants for one substep of length dt (host double -> float once).
Consts make_consts(const lf_params &p, float dt, int debug_flags) {
    Consts c;
    std::memset(&c, 0, sizeof(c));
    for (int a = 0; a < 3; ++a) {
        c.inv_cell[a] = p.inv_cell[a];
        c.half_box[a] = 0.5 * (double)p.box[a];
        c.g[a] = (double)p.gravity[a];
    }
    c.gx = (int)p.grid[0];
    c.gy = (int)p.grid[1];
    c.gz = (int)p.grid[2];
    c.ncells = p.grid[0] * p.grid[1] * p.grid[2];
    const double h = p.h, m = p.mass;
    const double h6 = std::pow(h, 6.0), h9 = std::pow(h, 9.0);
    const double C6 = 315.0 / (64.0 * kPi * h9);    // poly6 (fluidparams.cu:80)
    const double Cg6 = -945.0 / (32.0 * kPi * h9);  // gradPoly6 (fluidparams.cu:81)
    const double Cgs = -45.0 / (kPi * h6);          // gradSpiky (fluidparams.cu:84)
    c.h = p.h;
    c.h2 = p.h * p.h;  // FP32 product, the pair-predicate threshold (m_h2)
    c.inv_h = (float)(1.0 / h);
    c.eps = FLT_MIN;
    c.skin_q = (float)(2.2 * 4096.0 / h);  // the substep's skin is sized on the device from its largest move
    c.use_masks = (debug_flags & LF_DEBUG_SCAN) ? 0 : 1;
    const double h2f = (double)c.h2, h6f = h2f * h2f * h2f;
    c.inv_h2 = (float)(1.0 / h2f);
    c.mC6h6 = (float)(m * C6 * h6f);
    c.rho0 = p.rest_density;
    const double inv_rho0 = p.rest_density == 0.0f ? 1.0 : 1.0 / (double)p.rest_density;
    const double inv_gamma = p.gamma == 0.0f ? 1.0 : 1.0 / (double)p.gamma;
    c.inv_rho0 = (float)inv_rho0;
    c.eos_scale = (float)((double)p.gas_constant * (double)p.rest_density * inv_gamma);
    c.gamma = p.gamma;
    c.gamma_m1 = p.gamma - 1.0f;
    const auto small_int = [](float e) { return (e >= 1.0f && e <= 15.0f && e == std::floor(e)) ? (int)e : 0; };
    c.gamma_i = small_int(c.gamma);
    c.gamma_m1_i = small_int(c.gamma_m1);
    c.K = p.gas_constant;
    c.bd_rho0_d = (double)p.rest_density * (double)0.0039062f;  // fluidkernel.cu:65
    c.h_d = h;
    c.hh_d = h * h;
    c.mC6_d = m * C6;
    const double fragile_rho = (double)p.rest_density / 64.0;
    c.fragile_S = m * C6 > 0.0 ? (float)(fragile_rho / (m * C6 * h6f)) : 0.0f;
    c.fragile_ir = (float)(1.0 / fragile_rho);
    c.fragile_sr = (float)(2.0 * fragile_rho);
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
    // Normals feed only the surface-tension term; without it the force walk
    // computes them itself (same bits) and the separate pass is not needed.
    c.lazy_normal = (!c.has_tension && !(debug_flags & LF_DEBUG_NORMALS_PASS)) ? 1 : 0;
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

// ---------------------------------------------------------------------------
// lifetime

void Pool::release() {
    // Work still queued on the stream (setup copies of a failed init, the
    // last substeps or exports) may use any of the buffers below.
    if (stream_) (void)hipStreamSynchronize(stream_);
    stop_exporter();
    drop_graph();
    dev_free(MS_.w32);
    dev_free(MS_.w64);
    float4 **f4[] = {&P0_, &P1_, &V0_, &V1_, &V2_, &A0_, &A1_, &L0_, &L1_, &D_, &N_, &F_, &Q_};
    for (float4 **b : f4) dev_free(*b);
    dev_free(segtmp_);
    dev_free(Pr_);
    dev_free(soa_);
    uint32_t **u32[] = {&keys_, &ranks_, &slot_src_, &slot_id_, &counts_, &start_, &bsum_, &nsegs_, &flags_,
                        &moves_, &d_chunk_};
    for (uint32_t **b : u32) dev_free(*b);
    dev_free(segs_);
    dev_free(near_);
    dev_free(d_cs_);
    for (auto &r : prof_live_) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : ev_pool_) (void)hipEventDestroy(e);
    prof_live_.clear();
    ev_pool_.clear();
    for (int k = 0; k < kTableRing; ++k) {
        if (pin_cs_[k]) (void)hipHostFree(pin_cs_[k]);
        if (pin_chunk_[k]) (void)hipHostFree(pin_chunk_[k]);
        if (tables_ev_[k]) (void)hipEventDestroy(tables_ev_[k]);
        pin_cs_[k] = nullptr;
        pin_chunk_[k] = nullptr;
        tables_ev_[k] = nullptr;
        tables_ev_pending_[k] = false;
    }
    if (stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
}

Pool::~Pool() { release(); }

int Pool::init(uint32_t slot_cap, uint32_t bucket_cap, uint32_t scene_cap) {
    if (slot_cap == 0 || bucket_cap == 0 || scene_cap == 0) return fail(LF_ERR_INVALID, "pool: zero capacity");
    const uint64_t slots = ((uint64_t)slot_cap + kChunk - 1) / kChunk * kChunk;
    if (slots >= kMaxParticles)
        return fail(LF_ERR_INVALID, "pool: %llu particle slots exceed the limit of %u", (unsigned long long)slots,
                    kMaxParticles - 1u);
    if ((uint64_t)bucket_cap + 1 >= (1ull << 32)) return fail(LF_ERR_INVALID, "pool: too many buckets");
    nslots_ = (uint32_t)slots;
    nbuckets_ = bucket_cap;
    span_slots_ = nslots_;
    span_buckets_ = nbuckets_;
    ntiles_ = scan_tiles(nbuckets_);
    scenes_.assign(scene_cap, SceneRec{});
    h_cs_.assign(scene_cap, Consts{});
    h_chunk_.assign(nslots_ / kChunk, kNoScene);
    int rc = LF_OK;
    hipError_t e = hipGetDevice(&device_);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "pool: %s", hipGetErrorString(e));
    const size_t n = nslots_;
    float4 **f4[] = {&P0_, &P1_, &V0_, &V1_, &V2_, &A0_, &A1_, &L0_, &L1_, &D_, &N_, &Q_};
    for (float4 **b : f4)
        if (!rc) rc = dalloc(b, n);
    if (!rc) rc = dalloc(&segtmp_, n);
    if (!rc) rc = dalloc(&Pr_, n);
    if (!rc) rc = dalloc(&soa_, 3 * n);
    if (!rc) rc = dalloc(&MS_.w32, 9 * n + 2);  // +2: an 8-byte read at the last entry stays inside
    if (!rc) rc = dalloc(&MS_.w64, 9 * n);
    uint32_t **u32[] = {&keys_, &ranks_, &slot_src_, &slot_id_};
    for (uint32_t **b : u32)
        if (!rc) rc = dalloc(b, n);
    if (!rc) rc = dalloc(&counts_, (size_t)nbuckets_);
    if (!rc) rc = dalloc(&start_, (size_t)nbuckets_ + 1);
    if (!rc) rc = dalloc(&bsum_, (size_t)ntiles_ + 1);
    if (!rc) rc = dalloc(&nsegs_, 1);
    if (!rc) rc = dalloc(&segs_, n / (kRankDirectMax + 1) + 1);  // each segment holds > kRankDirectMax slots
    if (!rc) rc = dalloc(&flags_, (size_t)kSceneWords * scene_cap);
    if (!rc) rc = dalloc(&moves_, (size_t)kMoveSlots * kMoveStride * scene_cap);
    if (!rc) rc = dalloc(&near_, n / kNearBlock);
    if (!rc) rc = dalloc(&d_cs_, scene_cap);
    if (!rc) rc = dalloc(&d_chunk_, h_chunk_.size());
    for (int k = 0; k < kTableRing && !rc; ++k) {
        e = hipHostMalloc((void **)&pin_cs_[k], sizeof(Consts) * scene_cap, hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc((void **)&pin_chunk_[k], sizeof(uint32_t) * h_chunk_.size(), hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&tables_ev_[k], hipEventDisableTiming);
        if (e != hipSuccess) rc = fail(LF_ERR_ALLOC, "pool: pinned staging: %s", hipGetErrorString(e));
    }
    if (!rc) {
        // Counts start at zero; afterwards every scan clears them.
        e = hipMemsetAsync(counts_, 0, sizeof(uint32_t) * nbuckets_, stream_);
        if (e == hipSuccess) e = hipMemsetAsync(start_, 0, sizeof(uint32_t) * (nbuckets_ + 1), stream_);
        if (e == hipSuccess) e = hipMemsetAsync(nsegs_, 0, sizeof(uint32_t), stream_);
        if (e == hipSuccess) e = hipMemsetAsync(flags_, 0, sizeof(uint32_t) * kSceneWords * scene_cap, stream_);
        if (e == hipSuccess)
            e = hipMemsetAsync(moves_, 0, sizeof(uint32_t) * kMoveSlots * kMoveStride * scene_cap, stream_);
        if (e == hipSuccess) e = hipMemsetAsync(near_, 0, n / kNearBlock, stream_);
        if (e != hipSuccess) rc = fail(LF_ERR_HIP, "pool: %s", hipGetErrorString(e));
    }
    if (!rc) rc = push_tables();
    if (!rc) {
        e = hipStreamSynchronize(stream_);
        if (e != hipSuccess) rc = fail(LF_ERR_HIP, "pool: %s", hipGetErrorString(e));
    }
    if (rc) release();
    return rc;
}

// ---------------------------------------------------------------------------
// scene table

Consts Pool::make_scene_consts(uint32_t id) const {
    const SceneRec &s = scenes_[id];
    Consts c = make_consts(s.params, s.dt, debug_);
    c.first = s.first;
    c.count = s.n;
    c.cell_base = s.cell_base;
    c.scene = id;
    return c;
}

// Uploads the scene table (entries up to the highest live scene) and the
// chunk map when they changed: stream-ordered copies from pinned staging,
// which is reused once the previous copy from it has run.
int Pool::push_tables() {
    if (!tables_dirty_) return LF_OK;
    std::fill(h_chunk_.begin(), h_chunk_.end(), kNoScene);
    uint32_t top = 0;
    for (uint32_t id = 0; id < scenes_.size(); ++id) {
        const SceneRec &s = scenes_[id];
        if (!s.live) continue;
        h_cs_[id] = make_scene_consts(id);
        top = id + 1;
        if (s.active)
            for (uint32_t k = 0; k < s.nchunks; ++k) h_chunk_[s.first / kChunk + k] = id;
    }
    set_spans();
    const int k = table_slot_;
    table_slot_ = (table_slot_ + 1) % kTableRing;
    if (tables_ev_pending_[k]) LF_HIP(hipEventSynchronize(tables_ev_[k]));
    std::memcpy(pin_cs_[k], h_cs_.data(), sizeof(Consts) * top);
    std::memcpy(pin_chunk_[k], h_chunk_.data(), sizeof(uint32_t) * h_chunk_.size());
    if (top) LF_HIP(hipMemcpyAsync(d_cs_, pin_cs_[k], sizeof(Consts) * top, hipMemcpyHostToDevice, stream_));
    LF_HIP(hipMemcpyAsync(d_chunk_, pin_chunk_[k], sizeof(uint32_t) * h_chunk_.size(), hipMemcpyHostToDevice,
                          stream_));
    LF_HIP(hipEventRecord(tables_ev_[k], stream_));
    tables_ev_pending_[k] = true;
    tables_dirty_ = false;
    return LF_OK;
}

void Pool::set_spans() {
    const uint32_t gs = ((nslots_ / 16 + kChunk - 1) / kChunk) * kChunk, gb = (nbuckets_ + 15) / 16;
    uint64_t es = 0, eb = 0;
    for (const SceneRec &s : scenes_)
        if (s.live && s.active) {
            es = std::max<uint64_t>(es, s.first + (uint64_t)s.nchunks * kChunk);
            eb = std::max<uint64_t>(eb, s.cell_base + (uint64_t)s.ncells + 1);
        }
    es = std::max<uint64_t>(es, 1);
    eb = std::max<uint64_t>(eb, 1);
    span_slots_ = (uint32_t)std::min<uint64_t>(nslots_, (es + gs - 1) / gs * gs);
    span_buckets_ = (uint32_t)std::min<uint64_t>(nbuckets_, (eb + gb - 1) / gb * gb);
}

const SceneRec *Pool::scene(uint32_t id) const {
    if (id >= scenes_.size() || !scenes_[id].live) return nullptr;
    return &scenes_[id];
}

uint32_t Pool::live_scenes() const {
    uint32_t k = 0;
    for (const auto &s : scenes_) k += s.live ? 1u : 0u;
    return k;
}

void Pool::usage(uint32_t *slots_used, uint32_t *buckets_used) const {
    uint64_t a = 0, b = 0;
    for (const auto &s : scenes_)
        if (s.live) {
            a += (uint64_t)s.nchunks * kChunk;
            b += s.ncells + 1ull;
        }
    if (slots_used) *slots_used = (uint32_t)a;
    if (buckets_used) *buckets_used = (uint32_t)b;
}

// First fit over the chunk map (host copy of live scenes, active or not).
int Pool::find_slots(uint32_t nchunks, uint32_t *first) const {
    std::vector<std::pair<uint32_t, uint32_t>> used;  // [begin, end) in chunks
    for (const auto &s : scenes_)
        if (s.live) used.emplace_back(s.first / kChunk, s.first / kChunk + s.nchunks);
    std::sort(used.begin(), used.end());
    uint32_t at = 0;
    for (auto &u : used) {
        if (u.first - at >= nchunks) break;
        at = std::max(at, u.second);
    }
    if ((uint64_t)at + nchunks > h_chunk_.size()) return LF_ERR_FULL;
    *first = at * kChunk;
    return LF_OK;
}

int Pool::find_buckets(uint32_t nb, uint32_t *base) const {
    std::vector<std::pair<uint32_t, uint32_t>> used;
    for (const auto &s : scenes_)
        if (s.live) used.emplace_back(s.cell_base, s.cell_base + s.ncells + 1);
    std::sort(used.begin(), used.end());
    uint64_t at = 0;
    for (auto &u : used) {
        if (u.first - at >= nb) break;
        at = std::max<uint64_t>(at, u.second);
    }
    if (at + nb > nbuckets_) return LF_ERR_FULL;
    *base = (uint32_t)at;
    return LF_OK;
}

// Packs every live scene to the front of the slot and bucket spaces, keeping
// their order.  State moves through the scratch arrays (P1, V1, A1, L1, N),
// which hold nothing between substeps.
int Pool::compact() {
    std::vector<uint32_t> order;
    for (uint32_t id = 0; id < scenes_.size(); ++id)
        if (scenes_[id].live) order.push_back(id);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return scenes_[a].first < scenes_[b].first; });
    uint32_t at = 0;
    for (uint32_t id : order) {
        SceneRec &s = scenes_[id];
        if (s.first != at) {
            const uint32_t len = s.nchunks * kChunk;
            launch_move(stream_, s.first, s.first, len, P0_, V0_, A0_, L0_, D_, P1_, V1_, A1_, L1_, N_);
            launch_move(stream_, s.first, at, len, P1_, V1_, A1_, L1_, N_, P0_, V0_, A0_, L0_, D_);
            if (int rc = check_launch("compact")) return rc;
            s.first = at;
        }
        at += s.nchunks * kChunk;
    }
    std::sort(order.begin(), order.end(),
              [&](uint32_t a, uint32_t b) { return scenes_[a].cell_base < scenes_[b].cell_base; });
    uint32_t cb = 0;
    for (uint32_t id : order) {
        scenes_[id].cell_base = cb;  // counts are zero between substeps
        cb += scenes_[id].ncells + 1;
    }
    for (auto &s : scenes_) s.fields = false;
    tables_dirty_ = true;
    return LF_OK;
}

int Pool::add_scene(const lf_params &p, uint32_t n, float dt, uint32_t *id) {
    if (!id) return fail(LF_ERR_INVALID, "add_scene: null id");
    if (int rc = lf_params_check(&p)) return rc;
    if (n == 0) return fail(LF_ERR_INVALID, "add_scene: zero particles");
    if (!(std::isfinite(dt) && dt >= 0.0f)) return fail(LF_ERR_INVALID, "add_scene: bad dt");
    const uint64_t ncells = (uint64_t)p.grid[0] * p.grid[1] * p.grid[2];
    const uint32_t nchunks = (n + kChunk - 1) / kChunk;
    uint32_t sid = 0;
    while (sid < scenes_.size() && scenes_[sid].live) ++sid;
    if (sid == scenes_.size()) return fail(LF_ERR_FULL, "add_scene: all %zu scene entries in use", scenes_.size());
    if ((uint64_t)nchunks * kChunk > nslots_ || ncells + 1 > nbuckets_)
        return fail(LF_ERR_FULL, "add_scene: %u particles / %llu cells exceed the pool (%u slots, %u buckets)", n,
                    (unsigned long long)ncells, nslots_, nbuckets_);
    uint32_t first = 0, base = 0;
    int r1 = find_slots(nchunks, &first), r2 = find_buckets((uint32_t)ncells + 1, &base);
    if (r1 || r2) {
        uint32_t su = 0, bu = 0;
        usage(&su, &bu);
        if ((uint64_t)su + (uint64_t)nchunks * kChunk > nslots_ || (uint64_t)bu + ncells + 1 > nbuckets_)
            return fail(LF_ERR_FULL, "add_scene: pool full (%u of %u slots, %u of %u buckets in use)", su, nslots_,
                        bu, nbuckets_);
        if (int rc = compact()) return rc;
        r1 = find_slots(nchunks, &first);
        r2 = find_buckets((uint32_t)ncells + 1, &base);
        if (r1 || r2) return fail(LF_ERR_FULL, "add_scene: no contiguous room after compaction");
    }
    SceneRec s;
    s.live = true;
    s.active = true;
    s.params = p;
    s.n = n;
    s.nchunks = nchunks;
    s.first = first;
    s.ncells = (uint32_t)ncells;
    s.cell_base = base;
    s.dt = dt;
    s.identity = true;
    s.steps = 0;
    scenes_[sid] = s;
    tables_dirty_ = true;
    launch_init_state(stream_, first, n, P0_, V0_, A0_, L0_, D_);
    if (int rc = check_launch("add_scene")) {
        scenes_[sid].live = false;
        return rc;
    }
    *id = sid;
    return LF_OK;
}

int Pool::remove_scene(uint32_t id) {
    if (!scene(id)) return fail(LF_ERR_INVALID, "remove_scene: no scene %u", id);
    scenes_[id].live = false;
    scenes_[id].active = false;
    tables_dirty_ = true;
    return LF_OK;
}

int Pool::set_params(uint32_t id, const lf_params &p) {
    if (!scene(id)) return fail(LF_ERR_INVALID, "set_params: no scene %u", id);
    if (int rc = lf_params_check(&p)) return rc;
    const lf_params &o = scenes_[id].params;
    for (int a = 0; a < 3; ++a)
        if (p.grid[a] != o.grid[a] || p.inv_cell[a] != o.inv_cell[a] || p.box[a] != o.box[a])
            return fail(LF_ERR_INVALID, "set_params: the grid and box are fixed when the scene is created");
    scenes_[id].params = p;
    scenes_[id].fields = false;
    tables_dirty_ = true;
    drop_graph();  // the normals pass may come or go
    return LF_OK;
}

int Pool::set_dt(uint32_t id, float dt) {
    if (!scene(id)) return fail(LF_ERR_INVALID, "set_dt: no scene %u", id);
    if (!(std::isfinite(dt) && dt > 0.0f)) return fail(LF_ERR_INVALID, "set_dt: dt must be positive and finite");
    if (scenes_[id].dt != dt) {
        scenes_[id].dt = dt;
        tables_dirty_ = true;
    }
    return LF_OK;
}

int Pool::set_active(uint32_t id, bool on) {
    if (!scene(id)) return fail(LF_ERR_INVALID, "set_active: no scene %u", id);
    if (scenes_[id].active != on) {
        scenes_[id].active = on;
        tables_dirty_ = true;
        drop_graph();
    }
    return LF_OK;
}

// ---------------------------------------------------------------------------
// stepping

hipEvent_t Pool::take_event() {
    if (!ev_pool_.empty()) {
        hipEvent_t e = ev_pool_.back();
        ev_pool_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Runs `launch(L)`; when kernel class `kind` is profiled, L carries start and
// stop events that the launcher binds to its first and last dispatch.
template <class Fn>
void Pool::timed(int kind, Fn &&launch) {
    Launch L{stream_, nullptr, nullptr};
    if (prof_mask_ & (1u << kind)) {
        hipEvent_t a = take_event(), b = take_event();
        if (a && b) {
            L.a = a;
            L.b = b;
            launch(L);
            prof_live_.push_back(ProfRec{kind, a, b});
            return;
        }
        if (a) ev_pool_.push_back(a);
        if (b) ev_pool_.push_back(b);
    }
    launch(L);
}

bool Pool::any_separate_normals() const {
    for (uint32_t id = 0; id < scenes_.size(); ++id) {
        const SceneRec &s = scenes_[id];
        if (s.live && s.active && !make_consts(s.params, s.dt, debug_).lazy_normal) return true;
    }
    return false;
}

// Hash, scan, provisional scatter and stable gather of the state positions
// and velocities (P0, V0) into cell order (P1, V1); ranks_ then holds, per
// cell-order slot, the particle's state slot (the force pass reads A0 and L0
// there).
void Pool::enqueue_sort() {
    const PoolView pv = view();
    timed(LF_K_HASH, [&](const Launch &L) {
        launch_hash(L, pv, P0_, keys_, ranks_, counts_, flags_, moves_, nsegs_);
    });
    timed(LF_K_SCAN, [&](const Launch &L) { launch_scan(L, counts_, span_buckets_, bsum_, start_); });
    timed(LF_K_SCATTER, [&](const Launch &L) {
        launch_scatter(L, pv, P0_, keys_, ranks_, start_, slot_src_, slot_id_, segs_, nsegs_, segtmp_);
    });
    timed(LF_K_GATHER, [&](const Launch &L) {
        launch_gather(L, pv, slot_src_, slot_id_, keys_, start_, P0_, V0_, P1_, V1_, ranks_, soa_, moves_);
    });
}

// One substep of every active scene.  Buffer roles: state (P0,V0,A0,L0) ->
// sorted (P1,V1) -> integrated P0, A1, L1 (cell order; A0 and L0 read
// through ranks_) and pre-XSPH velocity V2 (density in w) -> XSPH velocity V0.  The new
// acceleration and low position words are in A1 and L1: advance_state()
// makes them the state.
void Pool::enqueue_substep() {
    const PoolView pv = view();
    enqueue_sort();
    timed(LF_K_DENSITY, [&](const Launch &L) {
        launch_density(L, pv, tune_.density, P1_, soa_, start_, D_, Q_, Pr_, MS_, flags_, moves_);
    });
    if (any_separate_normals())
        timed(LF_K_NORMALS, [&](const Launch &L) {
            launch_normals(L, pv, tune_.normals, tune_.flat_normals, false, Q_, start_, MS_, flags_, N_);
        });
    timed(LF_K_FORCES, [&](const Launch &L) {
        launch_forces(L, pv, tune_.forces, tune_.flat_forces, P1_, Q_, Pr_, V1_, A0_, L0_, ranks_, D_, N_, start_,
                      MS_, P0_, V2_, A1_, L1_, (debug_ & LF_DEBUG_FORCES) ? F_ : nullptr, flags_, moves_);
    });
    timed(LF_K_XSPH, [&](const Launch &L) {
        launch_xsph(L, pv, tune_.xsph, tune_.flat_xsph, P1_, P0_, V2_, L1_, start_, MS_, flags_, V0_);
    });
    for (const auto &r : paused_ranges()) launch_carry(stream_, r.first, r.second, A0_, L0_, A1_, L1_);
}

// Slot ranges of the live scenes that are paused (their A and L state is
// carried across each substep's buffer swap).
std::vector<std::pair<uint32_t, uint32_t>> Pool::paused_ranges() const {
    std::vector<std::pair<uint32_t, uint32_t>> r;
    for (const SceneRec &s : scenes_)
        if (s.live && !s.active && s.n) r.emplace_back(s.first, s.n);
    return r;
}

// After an enqueued substep: its acceleration and low position words (A1,
// L1) become the state; the old ones are scratch until the next substep.
void Pool::advance_state() {
    std::swap(A0_, A1_);
    std::swap(L0_, L1_);
    parity_ ^= 1;
}

// Launches of the executable graphs may still be queued: wait for them
// before destroying them.
void Pool::drop_graph() {
    if ((graph_exec_[0] || graph_exec_[1]) && stream_) (void)hipStreamSynchronize(stream_);
    for (int p = 0; p < 2; ++p) {
        if (graph_exec_[p]) (void)hipGraphExecDestroy(graph_exec_[p]);
        if (graph_[p]) (void)hipGraphDestroy(graph_[p]);
        graph_exec_[p] = nullptr;
        graph_[p] = nullptr;
    }
}

void Pool::set_graphs(bool on) {
    use_graphs_ = on;
    if (!on) drop_graph();
}

int Pool::run_substeps(uint32_t substeps) {
    if (int rc = push_tables()) return rc;
    if (!use_graphs_ || prof_mask_) {
        for (uint32_t k = 0; k < substeps; ++k) {
            enqueue_substep();
            advance_state();
        }
        return check_launch("step");
    }
    const std::pair<uint32_t, uint32_t> span{span_slots_, span_buckets_};
    if ((graph_exec_[0] || graph_exec_[1]) &&
        (graph_normals_ != any_separate_normals() || graph_paused_ != paused_ranges() || graph_span_ != span))
        drop_graph();
    for (uint32_t k = 0; k < substeps; ++k) {
        hipGraphExec_t &ge = graph_exec_[parity_];
        if (!ge) {
            // One substep captured once per pairing of the state and scratch
            // A/L buffers; the scene table it reads lives in device memory,
            // so later table changes need no new capture.
            LF_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            enqueue_substep();
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(stream_, &g);
            if (e != hipSuccess) return fail(LF_ERR_HIP, "step: graph capture failed: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                ge = nullptr;
                return fail(LF_ERR_HIP, "step: graph instantiation failed: %s", hipGetErrorString(e));
            }
            graph_[parity_] = g;
            graph_normals_ = any_separate_normals();
            graph_paused_ = paused_ranges();
            graph_span_ = span;
        }
        LF_HIP(hipGraphLaunch(ge, stream_));
        advance_state();
    }
    return LF_OK;
}

int Pool::step(uint32_t substeps) {
    if (substeps == 0) return fail(LF_ERR_INVALID, "step: zero substeps");
    bool any = false;
    for (const auto &s : scenes_)
        if (s.live && s.active) {
            if (!(s.dt > 0.0f)) return fail(LF_ERR_STATE, "step: an active scene has no positive dt");
            any = true;
        }
    if (!any) return LF_OK;
    if (int rc = run_substeps(substeps)) return rc;
    for (auto &s : scenes_) {
        s.fields = s.live && s.active;
        if (s.live && s.active) {
            s.steps += substeps;
            s.identity = false;
        }
    }
    return LF_OK;
}

int Pool::sync() {
    LF_HIP(hipStreamSynchronize(stream_));
    return LF_OK;
}

// ---------------------------------------------------------------------------
// per-scene data movement

// Brings one scene's state back to particle-id order (setup-time only).
int Pool::to_id_order(uint32_t id) {
    SceneRec &s = scenes_[id];
    if (s.identity) return LF_OK;
    launch_to_id_order(stream_, s.first, s.n, P0_, V0_, A0_, L0_, D_, P1_, V1_, A1_, L1_, N_);
    launch_move(stream_, s.first, s.first, s.n, P1_, V1_, A1_, L1_, N_, P0_, V0_, A0_, L0_, D_);
    if (int rc = check_launch("to_id_order")) return rc;
    s.identity = true;
    return LF_OK;
}

int Pool::upload(uint32_t id, uint32_t first, uint32_t count, const float *pos, const float *vel, const float *acc) {
    const SceneRec *s = scene(id);
    if (!s) return fail(LF_ERR_INVALID, "upload: no scene %u", id);
    if (!pos && count) return fail(LF_ERR_INVALID, "upload: null positions");
    if ((uint64_t)first + count > s->n)
        return fail(LF_ERR_INVALID, "upload: ids [%u, %llu) exceed the scene's %u particles", first,
                    (unsigned long long)first + count, s->n);
    if (count == 0) return LF_OK;
    if (int rc = to_id_order(id)) return rc;
    std::vector<float> stage(9 * (size_t)count, 0.0f);
    for (size_t k = 0; k < count; ++k)
        for (int a = 0; a < 3; ++a) {
            stage[9 * k + a] = pos[3 * k + a];
            if (vel) stage[9 * k + 3 + a] = vel[3 * k + a];
            if (acc) stage[9 * k + 6 + a] = acc[3 * k + a];
        }
    float *d_stage = nullptr;
    if (int rc = dalloc(&d_stage, stage.size())) return rc;
    hipError_t e = hipMemcpyAsync(d_stage, stage.data(), stage.size() * sizeof(float), hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) {
        launch_upload(stream_, d_stage, s->first, first, count, P0_, V0_, A0_, L0_);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    dev_free(d_stage);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "upload: %s", hipGetErrorString(e));
    return LF_OK;
}

int Pool::enqueue_upload(uint32_t id, const float *d_stage, uint32_t count) {
    const SceneRec *s = scene(id);
    if (!s || count > s->n) return fail(LF_ERR_INVALID, "enqueue_upload: bad scene %u or count", id);
    if (int rc = to_id_order(id)) return rc;
    launch_upload(stream_, d_stage, s->first, 0, count, P0_, V0_, A0_, L0_);
    return check_launch("enqueue_upload");
}

int Pool::enqueue_export(uint32_t id, float *d_out, uint32_t *d_ids, bool by_id) {
    if (!scene(id)) return fail(LF_ERR_INVALID, "enqueue_export: no scene %u", id);
    launch_export(stream_, make_scene_consts(id), P0_, V0_, A0_, D_, by_id, d_out, d_ids);
    return check_launch("enqueue_export");
}

int Pool::download(uint32_t id, int order, float *pos, float *col, float *vel, float *acc, uint32_t *ids) {
    const SceneRec *s = scene(id);
    if (!s) return fail(LF_ERR_INVALID, "download: no scene %u", id);
    if (order != LF_ORDER_CELL && order != LF_ORDER_ID) return fail(LF_ERR_INVALID, "download: bad order %d", order);
    const Consts c = make_scene_consts(id);
    const size_t n = s->n;
    float *d_out = nullptr;
    uint32_t *d_ids = nullptr;
    if (int rc = dalloc(&d_out, 12 * n)) return rc;
    if (int rc = dalloc(&d_ids, n)) {
        dev_free(d_out);
        return rc;
    }
    launch_export(stream_, c, P0_, V0_, A0_, D_, order == LF_ORDER_ID, d_out, d_ids);
    hipError_t e = hipGetLastError();
    std::vector<float> host(12 * n);
    std::vector<uint32_t> hid(n);
    if (e == hipSuccess) e = hipMemcpyAsync(host.data(), d_out, host.size() * sizeof(float), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(hid.data(), d_ids, hid.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    dev_free(d_out);
    dev_free(d_ids);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "download: %s", hipGetErrorString(e));
    for (size_t k = 0; k < n; ++k) {
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

// ---------------------------------------------------------------------------
// inspection

int Pool::set_debug(int flags) {
    if (flags & ~(LF_DEBUG_FORCES | LF_DEBUG_SCAN | LF_DEBUG_NORMALS_PASS))
        return fail(LF_ERR_INVALID, "set_debug: unknown flags %d", flags);
    if ((flags & LF_DEBUG_FORCES) && !F_)
        if (int rc = dalloc(&F_, 4 * (size_t)nslots_)) return rc;
    if (flags != debug_) {
        debug_ = flags;
        tables_dirty_ = true;
        drop_graph();
    }
    return LF_OK;
}

int Pool::get_field(uint32_t id, int field, void *dst, size_t bytes) {
    const SceneRec *s = scene(id);
    if (!s || !dst) return fail(LF_ERR_INVALID, "get_field: bad scene or null destination");
    LF_HIP(hipStreamSynchronize(stream_));
    const size_t n = s->n, f = s->first;
    const void *src = nullptr;
    size_t need = 0;
    switch (field) {
        case LF_FIELD_KEYS: src = keys_ + f; need = 4 * n; break;
        case LF_FIELD_START: src = start_ + s->cell_base; need = 4 * ((size_t)s->ncells + 2); break;
        case LF_FIELD_SORTED_IDS: need = 4 * n; break;
        case LF_FIELD_DENSITY: src = D_ + f; need = 16 * n; break;
        case LF_FIELD_NORMAL:
            // Normals feed only the surface-tension force; scenes without it
            // skip them while stepping.  They are computed here, for every
            // scene, from the last substep's data by the same pass.
            if (!s->fields)
                return fail(LF_ERR_STATE, "normals: no substep data for this scene (none since it was added, "
                                          "loaded, moved, re-parametrised or paused, or autotuning ran since)");
            launch_normals(Launch{stream_, nullptr, nullptr}, view(), tune_.normals, tune_.flat_normals, true,
                           Q_, start_, MS_, flags_, N_);
            if (int rc = check_launch("normals")) return rc;
            LF_HIP(hipStreamSynchronize(stream_));
            src = N_ + f;
            need = 16 * n;
            break;
        case LF_FIELD_FORCES:
            if (!F_) return fail(LF_ERR_STATE, "force components are only kept after set_debug(LF_DEBUG_FORCES)");
            src = F_ + 4 * f;
            need = 64 * n;
            break;
        case LF_FIELD_VEL_HALF: src = V2_ + f; need = 16 * n; break;
        case LF_FIELD_SORTED_POS: src = P1_ + f; need = 16 * n; break;
        default: return fail(LF_ERR_INVALID, "unknown field %d", field);
    }
    if (bytes != need) return fail(LF_ERR_INVALID, "field %d needs %zu bytes, got %zu", field, need, bytes);
    if (field == LF_FIELD_SORTED_IDS) {
        std::vector<float4> p(n);
        LF_HIP(hipMemcpy(p.data(), P1_ + f, 16 * n, hipMemcpyDeviceToHost));
        uint32_t *o = (uint32_t *)dst;
        for (size_t k = 0; k < n; ++k) std::memcpy(&o[k], &p[k].w, 4);
        return LF_OK;
    }
    LF_HIP(hipMemcpy(dst, src, need, hipMemcpyDeviceToHost));
    if (field == LF_FIELD_KEYS) {
        uint32_t *o = (uint32_t *)dst;
        for (size_t k = 0; k < n; ++k) {
            const uint32_t local = o[k] - s->cell_base;
            o[k] = local == s->ncells ? LF_NULL_HASH : local;
        }
    } else if (field == LF_FIELD_START) {
        uint32_t *o = (uint32_t *)dst;
        const uint32_t base = o[0];
        for (size_t k = 0; k < (size_t)s->ncells + 2; ++k) o[k] -= base;
    }
    return LF_OK;
}

int Pool::neighbours(uint32_t id, uint32_t *counts, uint32_t *list, uint64_t cap, uint64_t *total) {
    const SceneRec *s = scene(id);
    if (!s || !counts) return fail(LF_ERR_INVALID, "neighbours: bad scene or null counts");
    if (!s->active) return fail(LF_ERR_STATE, "neighbours: scene %u is not active", id);
    if (int rc = push_tables()) return rc;
    const Consts c = make_scene_consts(id);
    const uint32_t keep_mask = prof_mask_;
    prof_mask_ = 0;
    enqueue_sort();
    prof_mask_ = keep_mask;
    // The sort replaced the last substep's cell table and flags, which the
    // normals computed when read would use.
    for (auto &r : scenes_) r.fields = false;
    const size_t n = s->n;
    uint32_t *d_cnt = nullptr;
    if (int rc = dalloc(&d_cnt, n)) return rc;
    launch_nb_count(stream_, c, P1_, start_, d_cnt);
    std::vector<uint32_t> cnt_slot(n);
    std::vector<float4> p1(n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(cnt_slot.data(), d_cnt, 4 * n, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(p1.data(), P1_ + s->first, 16 * n, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    dev_free(d_cnt);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "neighbours: %s", hipGetErrorString(e));
    std::vector<uint32_t> slot_of(n);
    for (size_t k = 0; k < n; ++k) {
        uint32_t pid;
        std::memcpy(&pid, &p1[k].w, 4);
        slot_of[pid] = (uint32_t)k;
    }
    std::vector<uint64_t> off(n + 1, 0);
    for (size_t k = 0; k < n; ++k) off[k + 1] = off[k] + cnt_slot[k];
    const uint64_t pairs = off[n];
    if (total) *total = pairs;
    for (size_t pid = 0; pid < n; ++pid) counts[pid] = cnt_slot[slot_of[pid]];
    if (!list) return LF_OK;
    if (pairs > cap)
        return fail(LF_ERR_INVALID, "neighbours: %llu pairs exceed cap %llu", (unsigned long long)pairs,
                    (unsigned long long)cap);
    uint64_t *d_off = nullptr;
    uint32_t *d_list = nullptr;
    if (int rc = dalloc(&d_off, n + 1)) return rc;
    if (int rc = dalloc(&d_list, (size_t)std::max<uint64_t>(pairs, 1))) {
        dev_free(d_off);
        return rc;
    }
    std::vector<uint32_t> lst(pairs);
    e = hipMemcpyAsync(d_off, off.data(), 8 * (n + 1), hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) {
        launch_nb_fill(stream_, c, P1_, start_, d_off, d_list);
        e = hipGetLastError();
    }
    if (e == hipSuccess && pairs) e = hipMemcpyAsync(lst.data(), d_list, 4 * pairs, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    dev_free(d_off);
    dev_free(d_list);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "neighbours: %s", hipGetErrorString(e));
    uint64_t o = 0;
    for (size_t pid = 0; pid < n; ++pid) {
        const size_t k = slot_of[pid];
        std::copy(lst.begin() + off[k], lst.begin() + off[k + 1], list + o);
        std::sort(list + o, list + o + cnt_slot[k]);
        o += cnt_slot[k];
    }
    return LF_OK;
}

int Pool::profile_enable(uint32_t mask) {
    prof_mask_ = mask;
    return LF_OK;
}

int Pool::profile_read(uint64_t *launches, double *ms) {
    LF_HIP(hipStreamSynchronize(stream_));
    uint64_t cnt[LF_K_COUNT] = {0};
    double tot[LF_K_COUNT] = {0};
    for (auto &r : prof_live_) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess && r.kind >= 0 && r.kind < LF_K_COUNT) {
            cnt[r.kind]++;
            tot[r.kind] += t;
        }
        ev_pool_.push_back(r.a);
        ev_pool_.push_back(r.b);
    }
    prof_live_.clear();
    for (int k = 0; k < LF_K_COUNT; ++k) {
        if (launches) launches[k] = cnt[k];
        if (ms) ms[k] = tot[k];
    }
    return LF_OK;
}

int Pool::copy_probe(size_t bytes, int repeats, double *ms) {
    if (!ms || bytes == 0 || bytes % 16 || repeats <= 0) return fail(LF_ERR_INVALID, "copy_probe: bad argument");
    float4 *a = nullptr, *b = nullptr;
    if (int rc = dalloc(&a, bytes / 16)) return rc;
    if (int rc = dalloc(&b, bytes / 16)) {
        dev_free(a);
        return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemsetAsync(a, 0, bytes, stream_);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        const Launch plain{stream_, nullptr, nullptr};
        launch_copy(plain, a, b, bytes / 16);  // warm-up
        (void)hipEventRecord(e0, stream_);
        for (int r = 0; r < repeats; ++r) launch_copy(plain, a, b, bytes / 16);
        (void)hipEventRecord(e1, stream_);
        e = hipStreamSynchronize(stream_);
    }
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(a);
    dev_free(b);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "copy_probe: %s", hipGetErrorString(e));
    *ms = (double)t / repeats;
    return LF_OK;
}

int Pool::set_tuning(const Tuning &t) {
    if (!valid_block_size(t.density) || !valid_block_size(t.normals) || !valid_block_size(t.forces) ||
        !valid_block_size(t.xsph))
        return fail(LF_ERR_INVALID, "tuning: block sizes must be 128, 256 or 512");
    for (int f : {t.flat_normals, t.flat_forces, t.flat_xsph})
        if (f < 0 || f > 2)
            return fail(LF_ERR_INVALID,
                        "tuning: walk modes must be 0 (row by row), 1 (flattened) or 2 (flattened in waves with a "
                        "row of more than 32 slots)");
    tune_ = t;
    drop_graph();
    return LF_OK;
}

// Each candidate launch shape runs `reps` substeps from the same saved
// state, all four neighbour passes timed with events: first the three block
// sizes (walk modes as set), then, at the fastest block sizes, both walk
// modes of the normals, forces and XSPH passes.  The state (and every
// scene's step count and order flag) is restored afterwards, so tuning
// leaves the trajectories untouched.
int Pool::autotune(uint32_t reps, double ms_out[4][3]) {
    if (reps == 0) return fail(LF_ERR_INVALID, "autotune: zero reps");
    bool any = false;
    for (const auto &s : scenes_) any |= s.live && s.active && s.dt > 0.0f;
    if (!any) return fail(LF_ERR_STATE, "autotune: no active scene to time");
    if (int rc = push_tables()) return rc;
    const size_t n = nslots_;
    float4 *save = nullptr;
    if (int rc = dalloc(&save, 5 * n)) return rc;
    // The state arrays as they are now (substeps swap the A and L buffers).
    auto state = [&](int k) { return k == 0 ? P0_ : k == 1 ? V0_ : k == 2 ? A0_ : k == 3 ? L0_ : D_; };
    for (int k = 0; k < 5; ++k)
        LF_HIP(hipMemcpyAsync(save + k * n, state(k), 16 * n, hipMemcpyDeviceToDevice, stream_));
    std::vector<SceneRec> keep = scenes_;
    const Tuning before = tune_;
    const uint32_t keep_mask = prof_mask_;
    std::vector<ProfRec> keep_live;
    keep_live.swap(prof_live_);  // the caller's pending timings stay untouched
    prof_mask_ = 0;
    const int kinds[4] = {LF_K_DENSITY, LF_K_NORMALS, LF_K_FORCES, LF_K_XSPH};
    int rc = LF_OK;
    // Average ms per launch of each neighbour pass under launch shape t.
    auto time_shape = [&](const Tuning &t, double out[4]) {
        tune_ = t;
        for (int k = 0; k < 5 && !rc; ++k) {
            hipError_t e = hipMemcpyAsync(state(k), save + k * n, 16 * n, hipMemcpyDeviceToDevice, stream_);
            if (e != hipSuccess) rc = fail(LF_ERR_HIP, "autotune: %s", hipGetErrorString(e));
        }
        if (rc) return;
        enqueue_substep();  // warm-up, untimed
        advance_state();
        for (int k = 0; k < 4; ++k) prof_mask_ |= 1u << kinds[k];
        for (uint32_t r = 0; r < reps; ++r) {
            enqueue_substep();
            advance_state();
        }
        prof_mask_ = 0;
        if ((rc = check_launch("autotune"))) return;
        uint64_t cnt[LF_K_COUNT];
        double tot[LF_K_COUNT];
        if ((rc = profile_read(cnt, tot))) return;
        for (int k = 0; k < 4; ++k) out[k] = cnt[kinds[k]] ? tot[kinds[k]] / (double)cnt[kinds[k]] : 0.0;
    };
    const int sizes[3] = {128, 256, 512};
    double ms[4][3] = {};
    for (int b = 0; b < 3 && !rc; ++b) {
        Tuning t = before;
        t.density = t.normals = t.forces = t.xsph = sizes[b];
        double out[4] = {};
        time_shape(t, out);
        for (int k = 0; k < 4; ++k) ms[k][b] = out[k];
    }
    int best[4];
    for (int k = 0; k < 4; ++k) {
        best[k] = 1;
        for (int b = 0; b < 3; ++b)
            if (ms[k][b] > 0.0 && ms[k][b] < ms[k][best[k]]) best[k] = b;
    }
    Tuning chosen = before;
    chosen.density = sizes[best[0]];
    chosen.normals = sizes[best[1]];
    chosen.forces = sizes[best[2]];
    chosen.xsph = sizes[best[3]];
    double walk[3][4] = {};  // [row by row, flattened, flattened in waves with long rows][pass]
    for (int f = 0; f < 3 && !rc; ++f) {
        Tuning t = chosen;
        t.flat_normals = t.flat_forces = t.flat_xsph = f;
        time_shape(t, walk[f]);
    }
    auto pick = [&](int k) {  // flattened unless another mode was measurably faster
        int m = 1;
        for (int f : {2, 0})
            if (walk[f][k] > 0.0 && walk[f][k] < walk[m][k]) m = f;
        return m;
    };
    chosen.flat_normals = pick(1);
    chosen.flat_forces = pick(2);
    chosen.flat_xsph = pick(3);
    prof_mask_ = keep_mask;
    for (auto &r : prof_live_) {  // left over after an error
        ev_pool_.push_back(r.a);
        ev_pool_.push_back(r.b);
    }
    prof_live_.swap(keep_live);
    for (int k = 0; k < 5; ++k)
        (void)hipMemcpyAsync(state(k), save + k * n, 16 * n, hipMemcpyDeviceToDevice, stream_);
    hipError_t e = hipStreamSynchronize(stream_);
    dev_free(save);
    scenes_ = keep;
    for (auto &s : scenes_) s.fields = false;  // the working data are the timing runs
    if (!rc && e != hipSuccess) rc = fail(LF_ERR_HIP, "autotune: %s", hipGetErrorString(e));
    if (rc) {
        tune_ = before;
        drop_graph();
        return rc;
    }
    tune_ = chosen;
    drop_graph();
    if (ms_out)
        for (int k = 0; k < 4; ++k)
            for (int b = 0; b < 3; ++b) ms_out[k][b] = ms[k][b];
    return LF_OK;
}

}  // namespace lf
