// lf_pool.h -- host side of a pool of independent scenes that step together
// (internal to libfluid_hip.so).
//
// A pool owns one set of device arrays sized for `slot_cap` particle slots
// and `bucket_cap` cell buckets, plus a table of per-scene constants.  Scenes
// are admitted into free slot and bucket ranges (first fit, compacting when
// fragmented) and removed at any time; every substep is one launch per stage
// over the whole pool, either eager or replayed from a captured hipGraph.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "lf_abi.h"
#include "lf_device.h"
#include "lf_internal.h"
#include "lf_kernels.h"

namespace lf {

// Workgroup sizes of the four neighbour passes (128, 256 or 512 each).
// Launch shapes: workgroup size per neighbour pass, and whether the normals,
// forces and XSPH passes walk the neighbour masks as one flattened stream
// (1) or row by row (0).  Results do not depend on either.
struct Tuning {
    int density = 256, normals = 256, forces = 256, xsph = 256;
    int flat_normals = 1, flat_forces = 1, flat_xsph = 2;
};

struct SceneRec {
    bool live = false;
    bool active = false;       // stepped by step()
    lf_params params{};
    uint32_t n = 0;            // particles
    uint32_t nchunks = 0;      // slots [first, first + nchunks * kChunk)
    uint32_t first = 0;
    uint32_t ncells = 0;       // buckets [cell_base, cell_base + ncells + 1)
    uint32_t cell_base = 0;
    float dt = 0.0f;           // substep length
    bool identity = true;      // state arrays in particle-id order
    uint64_t steps = 0;
    // The last substep's working data (sorted positions, densities, masks,
    // cell table) are still in place: set by a substep that stepped this
    // scene; cleared by a substep without it, compaction, new parameters or
    // a loaded state.
    bool fields = false;
};

class Pool {
  public:
    Pool() = default;
    Pool(const Pool &) = delete;
    Pool &operator=(const Pool &) = delete;
    ~Pool();

    int init(uint32_t slot_cap, uint32_t bucket_cap, uint32_t scene_cap);

    // Scenes.  add_scene returns LF_ERR_FULL when the pool cannot hold it.
    int add_scene(const lf_params &p, uint32_t n, float dt, uint32_t *id);
    int remove_scene(uint32_t id);
    int set_params(uint32_t id, const lf_params &p);
    int set_dt(uint32_t id, float dt);
    int set_active(uint32_t id, bool on);
    const SceneRec *scene(uint32_t id) const;
    uint32_t scene_cap() const { return (uint32_t)scenes_.size(); }
    uint32_t slot_cap() const { return nslots_; }
    uint32_t bucket_cap() const { return nbuckets_; }
    uint32_t live_scenes() const;
    void usage(uint32_t *slots_used, uint32_t *buckets_used) const;

    // Data in / out of one scene (synchronous).
    int upload(uint32_t id, uint32_t first, uint32_t count, const float *pos, const float *vel, const float *acc);
    int download(uint32_t id, int order, float *pos, float *col, float *vel, float *acc, uint32_t *ids);
    // Stream-ordered variants for callers that stage through device memory
    // themselves: state of particles [0, count) from d_stage (9 floats each:
    // pos, vel, acc), and the scene's state in id order into d_out (12
    // floats each: pos, colour, vel, acc) and d_ids.
    hipStream_t stream() const { return stream_; }
    int enqueue_upload(uint32_t id, const float *d_stage, uint32_t count);
    int enqueue_export(uint32_t id, float *d_out, uint32_t *d_ids, bool by_id = true);

    // Exact state of one scene in id order, 12 floats per particle (position
    // high and low words, velocity, acceleration); synchronous.  load_state
    // also sets the scene's step count.
    int save_state(uint32_t id, std::vector<float> &out);
    int load_state(uint32_t id, const float *in, uint64_t steps);
    // Device-side setup of particle positions (ids [id0, ...)).
    int setup_jitter(uint32_t id, uint32_t id0, uint32_t nx, uint32_t ny, uint32_t nz, float dx, const float o[3],
                     float jitter, uint64_t seed);
    int setup_block(uint32_t id, uint32_t id0, uint32_t n, const float mn[3], const float mx[3], uint32_t *written,
                    float *mass, float *psize);
    // Asynchronous .geo export of one scene's current state (cell or id
    // order): a snapshot is taken on the stream, copied to host memory on a
    // second stream and written by a background thread while stepping goes
    // on.  export_wait waits for every pending file and reports the first
    // error since the last wait.
    int export_geo_async(uint32_t id, const char *filebase, uint32_t cnt, int order);
    int export_wait();

    // Enqueues `substeps` substeps of every active scene and returns.
    int step(uint32_t substeps);
    int sync();

    // Inspection of the last substep, tuning, timing.
    int set_debug(int flags);
    int get_field(uint32_t id, int field, void *dst, size_t bytes);
    int neighbours(uint32_t id, uint32_t *counts, uint32_t *list, uint64_t cap, uint64_t *total);
    int profile_enable(uint32_t mask);
    int profile_read(uint64_t *launches, double *ms);
    int copy_probe(size_t bytes, int repeats, double *ms);
    int set_tuning(const Tuning &t);
    const Tuning &tuning() const { return tune_; }
    // Times every block size of the neighbour passes on the pool's current
    // scenes (reps substeps each, state restored afterwards) and keeps the
    // fastest per pass.  ms[k][b]: average time of pass k (density, normals,
    // forces, xsph) at block size 128 << b; may be null.
    int autotune(uint32_t reps, double ms[4][3]);
    void set_graphs(bool on);

  private:
    struct Exporter;
    void stop_exporter();
    struct ProfRec {
        int kind;
        hipEvent_t a, b;
    };
    template <class T>
    int dalloc(T **p, size_t count);
    void release();
    Consts make_scene_consts(uint32_t id) const;
    int push_tables();
    int find_slots(uint32_t nchunks, uint32_t *first) const;
    int find_buckets(uint32_t nb, uint32_t *base) const;
    int compact();
    int to_id_order(uint32_t id);
    bool any_separate_normals() const;
    void enqueue_sort();
    void enqueue_substep();
    void advance_state();
    std::vector<std::pair<uint32_t, uint32_t>> paused_ranges() const;
    int run_substeps(uint32_t substeps);
    hipEvent_t take_event();
    template <class Fn>
    void timed(int kind, Fn &&launch);
    void drop_graph();
    PoolView view() const { return PoolView{d_cs_, d_chunk_, nslots_, span_slots_, near_}; }
    void set_spans();

    int device_ = 0;
    hipStream_t stream_ = nullptr;
    uint32_t nslots_ = 0, nbuckets_ = 0, ntiles_ = 0;
    // Slots and buckets the stepping launches cover, from the tables last
    // pushed: the end of the highest active scene's chunks and of its bucket
    // range, each rounded up to a sixteenth of the pool (so that a captured
    // substep stays valid while scenes come and go below that mark).
    uint32_t span_slots_ = 0, span_buckets_ = 0;
    int debug_ = 0;
    Tuning tune_{};
    bool use_graphs_ = true;

    // device arrays: state (P0,V0,A0,L0) and the substep's working arrays
    float4 *P0_ = nullptr, *P1_ = nullptr, *V0_ = nullptr, *V1_ = nullptr, *V2_ = nullptr;
    float4 *A0_ = nullptr, *A1_ = nullptr, *L0_ = nullptr, *L1_ = nullptr;
    float4 *D_ = nullptr, *N_ = nullptr, *F_ = nullptr, *Q_ = nullptr;
    uint2 *segtmp_ = nullptr;  // the segment sort's scratch, one entry per slot
    float *Pr_ = nullptr;
    float *soa_ = nullptr;  // sorted positions as three planes of nslots floats (x, y, z)
    RowMasks MS_{nullptr, nullptr};  // neighbour masks at the skin radius (density pass)
    uint32_t *keys_ = nullptr, *ranks_ = nullptr, *slot_src_ = nullptr, *slot_id_ = nullptr;
    uint32_t *counts_ = nullptr, *start_ = nullptr, *bsum_ = nullptr;
    uint4 *segs_ = nullptr;      // buckets of more than kRankDirectMax particles: (first slot, size, id bits)
    uint32_t *nsegs_ = nullptr;  // how many segs_ the substep listed
    uint32_t *flags_ = nullptr;  // kSceneWords per scene: the substep's flag bytes
    uint32_t *moves_ = nullptr;  // kMoveSlots * kMoveStride per scene: the substep's largest move
    uint8_t *near_ = nullptr;    // one byte per kNearBlock slots: near-isolated marks
    Consts *d_cs_ = nullptr;
    uint32_t *d_chunk_ = nullptr;

    // host tables
    std::vector<SceneRec> scenes_;
    std::vector<Consts> h_cs_;
    std::vector<uint32_t> h_chunk_;
    bool tables_dirty_ = true;
    // pinned staging of the tables, a ring of kTableRing copies, each reused
    // once the event of its last copy has completed
    static constexpr int kTableRing = 4;
    Consts *pin_cs_[kTableRing] = {};
    uint32_t *pin_chunk_[kTableRing] = {};
    hipEvent_t tables_ev_[kTableRing] = {};
    bool tables_ev_pending_[kTableRing] = {};
    int table_slot_ = 0;

    // captured substep, one per pairing of the state and scratch A/L buffers
    hipGraph_t graph_[2] = {};
    hipGraphExec_t graph_exec_[2] = {};
    int parity_ = 0;              // which pairing is current (toggled by advance_state)
    bool graph_normals_ = false;  // the captured substeps have the normals pass
    std::vector<std::pair<uint32_t, uint32_t>> graph_paused_;  // and carry these paused scenes' slots
    std::pair<uint32_t, uint32_t> graph_span_{0, 0};         // and cover these spans

    Exporter *exporter_ = nullptr;

    // profiling
    uint32_t prof_mask_ = 0;
    std::vector<ProfRec> prof_live_;
    std::vector<hipEvent_t> ev_pool_;
};

// Kernel constants of a scene with parameters p, substep dt and the given
// debug flags (placement fields left zero).
Consts make_consts(const lf_params &p, float dt, int debug_flags);

template <class T>
int Pool::dalloc(T **p, size_t count) {
    void *v = nullptr;
    const int rc = dev_alloc(&v, (count == 0 ? 1 : count) * sizeof(T));
    *p = static_cast<T *>(v);
    return rc;
}

}  // namespace lf
