// lf_scene.hip -- the device-facing half of the C ABI (include/lf_abi.h):
// lf_scene (one scene, the FluidSystem replacement) and lf_batch (many
// independent scenes stepped together), both thin layers over lf::Pool.
//
// lf_step / lf_batch_step enqueue work and return; lf_sync / lf_batch_sync
// wait and report deferred device errors.

#include <hip/hip_runtime.h>

#include <cmath>
#include <new>

#include "lf_internal.h"
#include "lf_io.h"
#include "lf_pool.h"

using lf::fail;

struct lf_scene {
    lf::Pool pool;
    uint32_t id = 0;
};

struct lf_batch {
    lf::Pool pool;
};

namespace {

uint64_t cells_of(const lf_params *p) { return (uint64_t)p->grid[0] * p->grid[1] * p->grid[2]; }

// Block sizes replace those of t; its walk modes stay.
int tuning_from(const int bs[4], lf::Tuning *t) {
    if (!bs) return fail(LF_ERR_INVALID, "null block sizes");
    t->density = bs[0];
    t->normals = bs[1];
    t->forces = bs[2];
    t->xsph = bs[3];
    return LF_OK;
}

int walks_from(const int flat[3], lf::Tuning *t) {
    if (!flat) return fail(LF_ERR_INVALID, "null walk modes");
    t->flat_normals = flat[0];
    t->flat_forces = flat[1];
    t->flat_xsph = flat[2];
    return LF_OK;
}

void walks_to(const lf::Tuning &t, int flat[3]) {
    flat[0] = t.flat_normals;
    flat[1] = t.flat_forces;
    flat[2] = t.flat_xsph;
}

void tuning_to(const lf::Tuning &t, int bs[4]) {
    bs[0] = t.density;
    bs[1] = t.normals;
    bs[2] = t.forces;
    bs[3] = t.xsph;
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

// ---- lf_scene ---------------------------------------------------------------

int lf_create(const lf_params *p, uint32_t num_particles, lf_scene **out) {
    if (!out) return fail(LF_ERR_INVALID, "lf_create: null output handle");
    *out = nullptr;
    if (int rc = lf_params_check(p)) return rc;
    if (num_particles == 0) return fail(LF_ERR_INVALID, "lf_create: zero particles");
    if (num_particles >= lf::kMaxParticles - lf::kChunk)
        return fail(LF_ERR_INVALID, "lf_create: %u particles exceed the limit of %u per scene", num_particles,
                    lf::kMaxParticles - lf::kChunk - 1u);
    const uint64_t nb = cells_of(p) + 1;
    if (nb >= (1ull << 32)) return fail(LF_ERR_INVALID, "lf_create: grid too large");
    lf_scene *s = new (std::nothrow) lf_scene();
    if (!s) return fail(LF_ERR_ALLOC, "lf_create: out of host memory");
    int rc = s->pool.init(num_particles, (uint32_t)nb, 1);
    if (!rc) rc = s->pool.add_scene(*p, num_particles, 0.0f, &s->id);
    if (!rc) rc = s->pool.sync();
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return LF_OK;
}

int lf_destroy(lf_scene *s) {
    delete s;
    return LF_OK;
}

int lf_num_particles(const lf_scene *s, uint32_t *n) {
    if (!s || !n) return fail(LF_ERR_INVALID, "null argument");
    *n = s->pool.scene(s->id)->n;
    return LF_OK;
}

int lf_set_params(lf_scene *s, const lf_params *p) {
    if (!s || !p) return fail(LF_ERR_INVALID, "null argument");
    return s->pool.set_params(s->id, *p);
}

int lf_get_params(const lf_scene *s, lf_params *p) {
    if (!s || !p) return fail(LF_ERR_INVALID, "null argument");
    *p = s->pool.scene(s->id)->params;
    return LF_OK;
}

int lf_upload(lf_scene *s, uint32_t first, uint32_t count, const float *pos, const float *vel, const float *acc) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.upload(s->id, first, count, pos, vel, acc);
}

int lf_download(lf_scene *s, int order, float *pos, float *col, float *vel, float *acc, uint32_t *ids) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.download(s->id, order, pos, col, vel, acc, ids);
}

int lf_step(lf_scene *s, float dt, uint32_t substeps) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    if (!(std::isfinite(dt) && dt > 0.0f)) return fail(LF_ERR_INVALID, "lf_step: dt must be positive and finite");
    if (substeps == 0) return fail(LF_ERR_INVALID, "lf_step: zero substeps");
    // part_dt = full_dt / substeps (fluidsystem.cu:295); exactly `substeps`
    // substeps run (the reference's float-accumulated loop, :301, can run one
    // fewer).
    if (int rc = s->pool.set_dt(s->id, dt / float(substeps))) return rc;
    return s->pool.step(substeps);
}

int lf_sync(lf_scene *s) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.sync();
}

int lf_step_count(const lf_scene *s, uint64_t *count) {
    if (!s || !count) return fail(LF_ERR_INVALID, "null argument");
    *count = s->pool.scene(s->id)->steps;
    return LF_OK;
}

int lf_set_debug(lf_scene *s, int flags) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.set_debug(flags);
}

int lf_get_field(lf_scene *s, int field, void *dst, size_t bytes) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.get_field(s->id, field, dst, bytes);
}

int lf_neighbours(lf_scene *s, uint32_t *counts, uint32_t *list, uint64_t cap, uint64_t *total) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.neighbours(s->id, counts, list, cap, total);
}

int lf_debug_flat_table(const uint64_t masks[9], const uint32_t first_slot[9], uint32_t *entry_mask,
                        uint32_t *entry_base, uint32_t *count) {
    if (!masks || !first_slot || !entry_mask || !entry_base || !count)
        return fail(LF_ERR_INVALID, "lf_debug_flat_table: null argument");
    *count = lf::flat_table(masks, first_slot, [&](uint32_t k, uint32_t m, uint32_t base) {
        entry_mask[k] = m;
        entry_base[k] = base;
    });
    return LF_OK;
}

int lf_debug_fail_alloc(int64_t nth) {
    lf::debug_fail_alloc(nth);
    return LF_OK;
}

int lf_debug_live_allocs(int64_t *count) {
    if (!count) return fail(LF_ERR_INVALID, "null count");
    *count = lf::debug_live_allocs();
    return LF_OK;
}

int lf_profile_enable(lf_scene *s, int mask) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.profile_enable((uint32_t)mask);
}

int lf_profile_read(lf_scene *s, uint64_t *launches, double *ms) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.profile_read(launches, ms);
}

int lf_copy_probe(lf_scene *s, size_t bytes, int repeats, double *ms) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.copy_probe(bytes, repeats, ms);
}

int lf_set_tuning(lf_scene *s, const int block_sizes[4]) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    lf::Tuning t = s->pool.tuning();
    if (int rc = tuning_from(block_sizes, &t)) return rc;
    return s->pool.set_tuning(t);
}

int lf_set_walks(lf_scene *s, const int flat[3]) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    lf::Tuning t = s->pool.tuning();
    if (int rc = walks_from(flat, &t)) return rc;
    return s->pool.set_tuning(t);
}

int lf_get_walks(const lf_scene *s, int flat[3]) {
    if (!s || !flat) return fail(LF_ERR_INVALID, "null argument");
    walks_to(s->pool.tuning(), flat);
    return LF_OK;
}

int lf_get_tuning(const lf_scene *s, int block_sizes[4]) {
    if (!s || !block_sizes) return fail(LF_ERR_INVALID, "null argument");
    tuning_to(s->pool.tuning(), block_sizes);
    return LF_OK;
}

int lf_autotune(lf_scene *s, uint32_t reps, double *ms) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    if (!(s->pool.scene(s->id)->dt > 0.0f))
        return fail(LF_ERR_STATE, "lf_autotune: step the scene once first (its substep length is not set)");
    return s->pool.autotune(reps, reinterpret_cast<double(*)[3]>(ms));
}

int lf_set_graphs(lf_scene *s, int on) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    s->pool.set_graphs(on != 0);
    return LF_OK;
}

// ---- formats either side of the path ----------------------------------------

int lf_checkpoint_save(lf_scene *s, const char *path) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    std::vector<float> v;
    if (int rc = s->pool.save_state(s->id, v)) return rc;
    const lf::SceneRec *r = s->pool.scene(s->id);
    return lf::write_checkpoint(path, r->params, r->n, r->steps, r->dt, v);
}

int lf_checkpoint_load(const char *path, lf_scene **out) {
    if (!out) return fail(LF_ERR_INVALID, "lf_checkpoint_load: null output handle");
    *out = nullptr;
    lf_params p;
    uint32_t n = 0;
    uint64_t steps = 0;
    float dt = 0.0f;
    std::vector<float> v;
    if (int rc = lf::read_checkpoint(path, &p, &n, &steps, &dt, v)) return rc;
    lf_scene *s = nullptr;
    if (int rc = lf_create(&p, n, &s)) return rc;
    int rc = s->pool.load_state(s->id, v.data(), steps);
    if (!rc && dt > 0.0f) rc = s->pool.set_dt(s->id, dt);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return LF_OK;
}

int lf_setup_jittered_lattice(lf_scene *s, uint32_t first, uint32_t nx, uint32_t ny, uint32_t nz, float dx,
                              const float origin[3], float jitter, uint64_t seed) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.setup_jitter(s->id, first, nx, ny, nz, dx, origin, jitter, seed);
}

int lf_setup_fluid_block(lf_scene *s, uint32_t first, uint32_t n, const float mn[3], const float mx[3],
                         uint32_t *written, float *mass, float *particle_size) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.setup_block(s->id, first, n, mn, mx, written, mass, particle_size);
}

int lf_export_geo_async(lf_scene *s, const char *filebase, uint32_t cnt, int order) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.export_geo_async(s->id, filebase, cnt, order);
}

int lf_export_wait(lf_scene *s) {
    if (!s) return fail(LF_ERR_INVALID, "null scene");
    return s->pool.export_wait();
}

// ---- lf_batch ---------------------------------------------------------------

int lf_batch_footprint(uint32_t max_particles, uint32_t max_cells, uint32_t max_scenes, uint64_t *bytes) {
    if (!bytes) return fail(LF_ERR_INVALID, "null argument");
    const uint64_t slots = ((uint64_t)max_particles + lf::kChunk - 1) / lf::kChunk * lf::kChunk;
    // 12 float4 arrays, the segment sort's 8-byte scratch, pressure, sorted
    // positions as three float planes, 9 x (32 + 64)-bit masks, four uint32
    // sort arrays per slot;
    // counts and starts per bucket; per scene its constants, flag word and
    // move table (kMoveSlots words one cache line apart).
    const uint64_t per_slot = 12 * 16 + 8 + 4 + 3 * 4 + 9 * (4 + 8) + 4 * 4;
    const uint64_t per_scene = sizeof(lf::Consts) + 4 * lf::kSceneWords + 4 * lf::kMoveSlots * lf::kMoveStride;
    // plus the chunk map, near-isolated marks, and the sort's segment list
    // (one 16-byte entry per kRankDirectMax + 1 slots) and its count
    *bytes = slots * per_slot + (uint64_t)max_cells * 8 + (uint64_t)max_scenes * per_scene + slots / lf::kChunk * 4 +
             slots / lf::kNearBlock + (slots / (lf::kRankDirectMax + 1) + 1) * 16 + 4;
    return LF_OK;
}

int lf_batch_create(uint32_t max_particles, uint32_t max_cells, uint32_t max_scenes, lf_batch **out) {
    if (!out) return fail(LF_ERR_INVALID, "lf_batch_create: null output handle");
    *out = nullptr;
    lf_batch *b = new (std::nothrow) lf_batch();
    if (!b) return fail(LF_ERR_ALLOC, "lf_batch_create: out of host memory");
    if (int rc = b->pool.init(max_particles, max_cells, max_scenes)) {
        delete b;
        return rc;
    }
    *out = b;
    return LF_OK;
}

int lf_batch_destroy(lf_batch *b) {
    delete b;
    return LF_OK;
}

int lf_batch_add(lf_batch *b, const lf_params *p, uint32_t num_particles, float dt, uint32_t *scene) {
    if (!b || !p) return fail(LF_ERR_INVALID, "null argument");
    return b->pool.add_scene(*p, num_particles, dt, scene);
}

int lf_batch_remove(lf_batch *b, uint32_t scene) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.remove_scene(scene);
}

int lf_batch_set_params(lf_batch *b, uint32_t scene, const lf_params *p) {
    if (!b || !p) return fail(LF_ERR_INVALID, "null argument");
    return b->pool.set_params(scene, *p);
}

int lf_batch_set_dt(lf_batch *b, uint32_t scene, float dt) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.set_dt(scene, dt);
}

int lf_batch_set_active(lf_batch *b, uint32_t scene, int active) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.set_active(scene, active != 0);
}

int lf_batch_upload(lf_batch *b, uint32_t scene, uint32_t first, uint32_t count, const float *pos, const float *vel,
                    const float *acc) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.upload(scene, first, count, pos, vel, acc);
}

int lf_batch_download(lf_batch *b, uint32_t scene, int order, float *pos, float *colour, float *vel, float *acc,
                      uint32_t *ids) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.download(scene, order, pos, colour, vel, acc, ids);
}

int lf_batch_step(lf_batch *b, uint32_t substeps) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.step(substeps);
}

int lf_batch_sync(lf_batch *b) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.sync();
}

int lf_batch_scene_info(const lf_batch *b, uint32_t scene, uint32_t *num_particles, uint64_t *steps, int *active) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    const lf::SceneRec *s = b->pool.scene(scene);
    if (!s) return fail(LF_ERR_INVALID, "no scene %u", scene);
    if (num_particles) *num_particles = s->n;
    if (steps) *steps = s->steps;
    if (active) *active = s->active ? 1 : 0;
    return LF_OK;
}

int lf_batch_usage(const lf_batch *b, uint32_t *scenes, uint32_t *slots_used, uint32_t *slot_cap,
                   uint32_t *buckets_used, uint32_t *bucket_cap) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    if (scenes) *scenes = b->pool.live_scenes();
    b->pool.usage(slots_used, buckets_used);
    if (slot_cap) *slot_cap = b->pool.slot_cap();
    if (bucket_cap) *bucket_cap = b->pool.bucket_cap();
    return LF_OK;
}

int lf_batch_set_tuning(lf_batch *b, const int block_sizes[4]) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    lf::Tuning t = b->pool.tuning();
    if (int rc = tuning_from(block_sizes, &t)) return rc;
    return b->pool.set_tuning(t);
}

int lf_batch_set_walks(lf_batch *b, const int flat[3]) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    lf::Tuning t = b->pool.tuning();
    if (int rc = walks_from(flat, &t)) return rc;
    return b->pool.set_tuning(t);
}

int lf_batch_get_walks(const lf_batch *b, int flat[3]) {
    if (!b || !flat) return fail(LF_ERR_INVALID, "null argument");
    walks_to(b->pool.tuning(), flat);
    return LF_OK;
}

int lf_batch_get_tuning(const lf_batch *b, int block_sizes[4]) {
    if (!b || !block_sizes) return fail(LF_ERR_INVALID, "null argument");
    tuning_to(b->pool.tuning(), block_sizes);
    return LF_OK;
}

int lf_batch_autotune(lf_batch *b, uint32_t reps, double *ms) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.autotune(reps, reinterpret_cast<double(*)[3]>(ms));
}

int lf_batch_set_graphs(lf_batch *b, int on) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    b->pool.set_graphs(on != 0);
    return LF_OK;
}

int lf_batch_checkpoint_save(lf_batch *b, uint32_t scene, const char *path) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    std::vector<float> v;
    if (int rc = b->pool.save_state(scene, v)) return rc;
    const lf::SceneRec *r = b->pool.scene(scene);
    return lf::write_checkpoint(path, r->params, r->n, r->steps, r->dt, v);
}

int lf_batch_checkpoint_load(lf_batch *b, const char *path, uint32_t *scene) {
    if (!b || !scene) return fail(LF_ERR_INVALID, "null argument");
    lf_params p;
    uint32_t n = 0;
    uint64_t steps = 0;
    float dt = 0.0f;
    std::vector<float> v;
    if (int rc = lf::read_checkpoint(path, &p, &n, &steps, &dt, v)) return rc;
    uint32_t id = 0;
    if (int rc = b->pool.add_scene(p, n, dt, &id)) return rc;
    if (int rc = b->pool.load_state(id, v.data(), steps)) {
        (void)b->pool.remove_scene(id);
        return rc;
    }
    *scene = id;
    return LF_OK;
}

int lf_batch_setup_jittered_lattice(lf_batch *b, uint32_t scene, uint32_t first, uint32_t nx, uint32_t ny,
                                    uint32_t nz, float dx, const float origin[3], float jitter, uint64_t seed) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.setup_jitter(scene, first, nx, ny, nz, dx, origin, jitter, seed);
}

int lf_batch_setup_fluid_block(lf_batch *b, uint32_t scene, uint32_t first, uint32_t n, const float mn[3],
                               const float mx[3], uint32_t *written, float *mass, float *particle_size) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.setup_block(scene, first, n, mn, mx, written, mass, particle_size);
}

int lf_batch_export_geo_async(lf_batch *b, uint32_t scene, const char *filebase, uint32_t cnt, int order) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.export_geo_async(scene, filebase, cnt, order);
}

int lf_batch_export_wait(lf_batch *b) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.export_wait();
}

int lf_batch_profile_enable(lf_batch *b, int mask) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.profile_enable((uint32_t)mask);
}

int lf_batch_profile_read(lf_batch *b, uint64_t *launches, double *ms) {
    if (!b) return fail(LF_ERR_INVALID, "null batch");
    return b->pool.profile_read(launches, ms);
}

}  // extern "C"
