// lf_serve.hip -- continuous-batching server over a scene pool (host code;
// C ABI lf_server_* in include/lf_abi.h).
//
// Jobs (a scene plus a number of substeps) are queued by lf_server_submit and
// run by lf_server_tick: each tick admits queued jobs into free pool memory
// (arrival order; with backfill, later jobs may pass one that does not fit
// yet), advances every running job by one substep in one batched step, and
// retires the jobs that reached their step count, freeing their pool memory
// for the queue.  A job's result is bitwise the one it would have alone,
// whatever else runs beside it.
//
// Everything a tick does on the device is ordered on the pool's one stream
// (admission uploads, the step, retirement exports), so a tick never waits
// for the device: a retired job's slots can be handed to a new job in the
// same tick because the new job's upload runs after the old job's export.
// Host staging is pinned memory, reused once the copy that last used it has
// run; at most run_ahead steps are queued ahead of the device.  Per-step
// latency is the device time of each batched step (events around it).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <new>
#include <unordered_map>
#include <vector>

#include "lf_internal.h"
#include "lf_pool.h"

using lf::fail;

#define LF_HIP(expr)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return lf::fail(LF_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

double quantile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t lo = (size_t)pos, hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

// Pinned host buffer with the event of the last copy that used it.  busy: a
// copy from or into it may still be in flight; held: it holds an exported
// state not yet moved out (until finish()).
struct Pinned {
    float *p = nullptr;
    size_t cap = 0;  // floats
    hipEvent_t ev = nullptr;
    bool busy = false;
    bool held = false;
};

enum { kRetiring = 3 };  // internal state: export queued, copy not yet complete

struct Job {
    uint64_t id = 0;
    int state = LF_JOB_QUEUED;
    lf_params params{};
    uint32_t n = 0, steps = 0, done = 0;
    float dt = 0.0f;
    uint32_t scene = 0;
    int pin = -1;                               // pinned buffer of the export in flight
    std::vector<float> pos, vel;                // initial state until admitted
    std::vector<float> rpos, rvel, racc, rcol;  // final state once done
    Clock::time_point t_submit, t_admit, t_done;
};

struct StepEv {
    hipEvent_t a, b;
};

}  // namespace

struct lf_server {
    lf::Pool pool;
    int policy = LF_SERVE_BACKFILL;
    size_t run_ahead = 2;  // batched steps queued ahead of the device at most (0: every tick waits)
    uint64_t next_id = 1;
    std::unordered_map<uint64_t, Job> jobs;
    std::deque<uint64_t> queue;       // queued job ids, arrival order
    std::vector<uint64_t> running;    // admitted job ids
    std::vector<uint64_t> retiring;   // exports in flight
    std::deque<uint64_t> unreported;  // done, not yet listed by lf_server_take_finished
    // staging
    std::vector<Pinned> pins;
    float *d_in = nullptr, *d_out = nullptr;
    uint32_t *d_ids = nullptr;
    size_t d_in_cap = 0, d_out_cap = 0;  // particles
    std::deque<StepEv> steps_live;
    std::vector<hipEvent_t> ev_free;
    // statistics
    uint64_t submitted = 0, admitted = 0, finished = 0, ticks = 0, step_ticks = 0, particle_steps = 0;
    double busy_ms = 0.0, occupancy_sum = 0.0;
    uint32_t resident = 0;
    std::vector<double> step_ms, job_ms, wait_ms;

    ~lf_server() {
        if (pool.stream()) (void)hipStreamSynchronize(pool.stream());
        for (auto &p : pins) {
            if (p.p) (void)hipHostFree(p.p);
            if (p.ev) (void)hipEventDestroy(p.ev);
        }
        for (auto &s : steps_live) {
            (void)hipEventDestroy(s.a);
            (void)hipEventDestroy(s.b);
        }
        for (auto e : ev_free) (void)hipEventDestroy(e);
        lf::dev_free(d_in);
        lf::dev_free(d_out);
        lf::dev_free(d_ids);
    }
};

namespace {

int take_event(lf_server *s, hipEvent_t *e) {
    if (!s->ev_free.empty()) {
        *e = s->ev_free.back();
        s->ev_free.pop_back();
        return LF_OK;
    }
    LF_HIP(hipEventCreate(e));
    return LF_OK;
}

// A pinned buffer of at least `floats` floats that no copy in flight uses
// and that holds no export waiting to be moved out.
int acquire_pinned(lf_server *s, size_t floats, int *idx) {
    int best = -1;
    for (size_t k = 0; k < s->pins.size(); ++k) {
        Pinned &p = s->pins[k];
        if (p.busy && p.ev && hipEventQuery(p.ev) == hipSuccess) p.busy = false;
        if (!p.busy && !p.held && p.cap >= floats && (best < 0 || p.cap < s->pins[best].cap)) best = (int)k;
    }
    if (best < 0) {
        Pinned p;
        p.cap = std::max<size_t>(floats, 1 << 16);
        LF_HIP(hipHostMalloc((void **)&p.p, p.cap * sizeof(float), hipHostMallocDefault));
        hipError_t e = hipEventCreateWithFlags(&p.ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipHostFree(p.p);
            return fail(LF_ERR_HIP, "pinned staging: %s", hipGetErrorString(e));
        }
        s->pins.push_back(p);
        best = (int)s->pins.size() - 1;
    }
    *idx = best;
    return LF_OK;
}

// Device staging large enough for n particles (grows geometrically; a
// reallocation waits for the device).
int ensure_device(lf_server *s, uint32_t n, bool out) {
    size_t &cap = out ? s->d_out_cap : s->d_in_cap;
    if (n <= cap) return LF_OK;
    LF_HIP(hipStreamSynchronize(s->pool.stream()));
    const size_t want = std::max<size_t>(n, 2 * cap);
    if (out) {
        lf::dev_free(s->d_out);
        lf::dev_free(s->d_ids);
        cap = 0;
        if (int rc = lf::dev_alloc((void **)&s->d_out, want * 12 * sizeof(float))) return rc;
        if (int rc = lf::dev_alloc((void **)&s->d_ids, want * sizeof(uint32_t))) return rc;
    } else {
        lf::dev_free(s->d_in);
        cap = 0;
        if (int rc = lf::dev_alloc((void **)&s->d_in, want * 9 * sizeof(float))) return rc;
    }
    cap = want;
    return LF_OK;
}

int admit(lf_server *s, Job &j) {
    uint32_t scene = 0;
    const int rc = s->pool.add_scene(j.params, j.n, j.dt, &scene);
    if (rc) return rc;
    int pin = -1;
    int r2 = ensure_device(s, j.n, false);
    if (!r2) r2 = acquire_pinned(s, 9 * (size_t)j.n, &pin);
    if (!r2) {
        float *h = s->pins[pin].p;
        for (size_t k = 0; k < j.n; ++k) {
            for (int a = 0; a < 3; ++a) {
                h[9 * k + a] = j.pos[3 * k + a];
                h[9 * k + 3 + a] = j.vel.empty() ? 0.0f : j.vel[3 * k + a];
                h[9 * k + 6 + a] = 0.0f;
            }
        }
        hipError_t e = hipMemcpyAsync(s->d_in, h, 9 * sizeof(float) * j.n, hipMemcpyHostToDevice, s->pool.stream());
        if (e == hipSuccess) e = hipEventRecord(s->pins[pin].ev, s->pool.stream());
        if (e != hipSuccess) r2 = fail(LF_ERR_HIP, "admit: %s", hipGetErrorString(e));
        s->pins[pin].busy = true;
    }
    if (!r2) r2 = s->pool.enqueue_upload(scene, s->d_in, j.n);
    if (r2) {
        (void)s->pool.remove_scene(scene);
        return r2;
    }
    j.scene = scene;
    j.state = LF_JOB_RUNNING;
    j.t_admit = Clock::now();
    std::vector<float>().swap(j.pos);
    std::vector<float>().swap(j.vel);
    s->running.push_back(j.id);
    s->resident += j.n;
    s->admitted++;
    s->wait_ms.push_back(ms_since(j.t_submit, j.t_admit));
    return LF_OK;
}

// Queues the export of a job that reached its step count and frees its pool
// memory (later uploads into it are ordered after this export).
int retire(lf_server *s, Job &j) {
    int pin = -1;
    if (int rc = ensure_device(s, j.n, true)) return rc;
    if (int rc = acquire_pinned(s, 12 * (size_t)j.n, &pin)) return rc;
    if (int rc = s->pool.enqueue_export(j.scene, s->d_out, s->d_ids)) return rc;
    LF_HIP(hipMemcpyAsync(s->pins[pin].p, s->d_out, 12 * sizeof(float) * j.n, hipMemcpyDeviceToHost,
                          s->pool.stream()));
    LF_HIP(hipEventRecord(s->pins[pin].ev, s->pool.stream()));
    s->pins[pin].busy = true;
    s->pins[pin].held = true;
    if (int rc = s->pool.remove_scene(j.scene)) return rc;
    j.pin = pin;
    j.state = kRetiring;
    s->resident -= j.n;
    s->retiring.push_back(j.id);
    return LF_OK;
}

// Moves the exported state of a retiring job out of its pinned buffer once
// the copy has run (waiting for it when `wait`).  Returns true when done.
bool finish(lf_server *s, Job &j, bool wait) {
    Pinned &p = s->pins[j.pin];
    if (wait)
        (void)hipEventSynchronize(p.ev);
    else if (hipEventQuery(p.ev) != hipSuccess)
        return false;
    const size_t n = j.n;
    j.rpos.resize(3 * n);
    j.rcol.resize(3 * n);
    j.rvel.resize(3 * n);
    j.racc.resize(3 * n);
    for (size_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) {
            j.rpos[3 * k + a] = p.p[12 * k + a];
            j.rcol[3 * k + a] = p.p[12 * k + 3 + a];
            j.rvel[3 * k + a] = p.p[12 * k + 6 + a];
            j.racc[3 * k + a] = p.p[12 * k + 9 + a];
        }
    p.busy = false;
    p.held = false;
    j.pin = -1;
    j.state = LF_JOB_DONE;
    j.t_done = Clock::now();
    s->finished++;
    s->job_ms.push_back(ms_since(j.t_submit, j.t_done));
    s->unreported.push_back(j.id);
    return true;
}

void collect(lf_server *s, bool wait) {
    std::vector<uint64_t> still;
    for (uint64_t id : s->retiring)
        if (!finish(s, s->jobs.at(id), wait)) still.push_back(id);
    s->retiring.swap(still);
}

// Harvests the device times of completed steps (in order).
void harvest_steps(lf_server *s, bool wait) {
    while (!s->steps_live.empty()) {
        StepEv &se = s->steps_live.front();
        if (wait)
            (void)hipEventSynchronize(se.b);
        else if (hipEventQuery(se.b) != hipSuccess)
            break;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, se.a, se.b) == hipSuccess) s->step_ms.push_back(ms);
        s->ev_free.push_back(se.a);
        s->ev_free.push_back(se.b);
        s->steps_live.pop_front();
    }
}

}  // namespace

extern "C" {

int lf_server_create(uint32_t max_particles, uint32_t max_cells, uint32_t max_scenes, int policy, lf_server **out) {
    if (!out) return fail(LF_ERR_INVALID, "lf_server_create: null output handle");
    *out = nullptr;
    if (policy != LF_SERVE_FIFO && policy != LF_SERVE_BACKFILL)
        return fail(LF_ERR_INVALID, "lf_server_create: unknown policy %d", policy);
    lf_server *s = new (std::nothrow) lf_server();
    if (!s) return fail(LF_ERR_ALLOC, "lf_server_create: out of host memory");
    if (int rc = s->pool.init(max_particles, max_cells, max_scenes)) {
        delete s;
        return rc;
    }
    s->policy = policy;
    *out = s;
    return LF_OK;
}

int lf_server_destroy(lf_server *s) {
    delete s;
    return LF_OK;
}

int lf_server_submit(lf_server *s, const lf_params *p, uint32_t n, const float *pos_xyz, const float *vel_xyz,
                     float dt, uint32_t steps, uint64_t *job) {
    if (!s || !p || !pos_xyz || !job) return fail(LF_ERR_INVALID, "lf_server_submit: null argument");
    if (int rc = lf_params_check(p)) return rc;
    if (n == 0 || steps == 0) return fail(LF_ERR_INVALID, "lf_server_submit: zero particles or steps");
    if (!(dt > 0.0f)) return fail(LF_ERR_INVALID, "lf_server_submit: dt must be positive");
    const uint64_t slots = ((uint64_t)n + lf::kChunk - 1) / lf::kChunk * lf::kChunk;
    const uint64_t cells = (uint64_t)p->grid[0] * p->grid[1] * p->grid[2] + 1;
    if (slots > s->pool.slot_cap() || cells > s->pool.bucket_cap())
        return fail(LF_ERR_FULL, "lf_server_submit: a job of %u particles / %llu cells can never fit the server", n,
                    (unsigned long long)cells);
    Job j;
    j.id = s->next_id++;
    j.params = *p;
    j.n = n;
    j.steps = steps;
    j.dt = dt;
    j.pos.assign(pos_xyz, pos_xyz + 3 * (size_t)n);
    if (vel_xyz) j.vel.assign(vel_xyz, vel_xyz + 3 * (size_t)n);
    j.t_submit = Clock::now();
    *job = j.id;
    s->queue.push_back(j.id);
    s->jobs.emplace(j.id, std::move(j));
    s->submitted++;
    return LF_OK;
}

int lf_server_tick(lf_server *s, uint32_t *finished) {
    if (!s) return fail(LF_ERR_INVALID, "null server");
    const Clock::time_point t0 = Clock::now();
    const uint64_t fin0 = s->finished;
    harvest_steps(s, false);
    collect(s, false);
    // Admission, arrival order.
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Job &j = s->jobs.at(*it);
        const int rc = admit(s, j);
        if (rc == LF_OK) {
            it = s->queue.erase(it);
            continue;
        }
        if (rc != LF_ERR_FULL) return rc;
        lf::clear_error();
        if (s->policy == LF_SERVE_FIFO) break;
        ++it;
    }
    if (!s->running.empty()) {
        if (s->run_ahead && s->steps_live.size() >= s->run_ahead) {  // bound the run-ahead
            LF_HIP(hipEventSynchronize(s->steps_live[s->steps_live.size() - s->run_ahead].b));
            harvest_steps(s, false);
        }
        StepEv se{};
        if (int rc = take_event(s, &se.a)) return rc;
        if (int rc = take_event(s, &se.b)) return rc;
        LF_HIP(hipEventRecord(se.a, s->pool.stream()));
        if (int rc = s->pool.step(1)) return rc;
        LF_HIP(hipEventRecord(se.b, s->pool.stream()));
        s->steps_live.push_back(se);
        s->particle_steps += s->resident;
        s->occupancy_sum += (double)s->resident;
        std::vector<uint64_t> still;
        still.reserve(s->running.size());
        for (uint64_t id : s->running) {
            Job &j = s->jobs.at(id);
            if (++j.done == j.steps) {
                if (int rc = retire(s, j)) return rc;
            } else {
                still.push_back(id);
            }
        }
        s->running.swap(still);
        s->step_ticks++;
        if (!s->run_ahead) {
            harvest_steps(s, true);
            collect(s, true);
        }
    }
    if (s->running.empty() && s->queue.empty()) {  // nothing left to overlap with
        harvest_steps(s, true);
        collect(s, true);
    }
    s->ticks++;
    s->busy_ms += ms_since(t0, Clock::now());
    if (finished) *finished = (uint32_t)(s->finished - fin0);
    return LF_OK;
}

int lf_server_configure(lf_server *s, uint32_t run_ahead, int graphs) {
    if (!s) return fail(LF_ERR_INVALID, "null server");
    if (run_ahead > 64) return fail(LF_ERR_INVALID, "lf_server_configure: run_ahead above 64");
    s->run_ahead = run_ahead;
    s->pool.set_graphs(graphs != 0);
    return LF_OK;
}

int lf_server_take_finished(lf_server *s, uint64_t *jobs, uint32_t cap, uint32_t *count) {
    if (!s || !count || (!jobs && cap)) return fail(LF_ERR_INVALID, "null argument");
    uint32_t k = 0;
    while (k < cap && !s->unreported.empty()) {
        jobs[k++] = s->unreported.front();
        s->unreported.pop_front();
    }
    *count = k;
    return LF_OK;
}

int lf_server_job_status(const lf_server *s, uint64_t job, int *state, uint32_t *steps_done) {
    if (!s) return fail(LF_ERR_INVALID, "null server");
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return fail(LF_ERR_INVALID, "no job %llu", (unsigned long long)job);
    if (state) *state = it->second.state == kRetiring ? LF_JOB_RUNNING : it->second.state;
    if (steps_done) *steps_done = it->second.done;
    return LF_OK;
}

int lf_server_job_result(lf_server *s, uint64_t job, float *pos_xyz, float *colour_rgb, float *vel_xyz,
                         float *acc_xyz, double *latency_ms) {
    if (!s) return fail(LF_ERR_INVALID, "null server");
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return fail(LF_ERR_INVALID, "no job %llu", (unsigned long long)job);
    Job &j = it->second;
    if (j.state == kRetiring) {
        (void)finish(s, j, true);
        s->retiring.erase(std::find(s->retiring.begin(), s->retiring.end(), job));
    }
    if (j.state != LF_JOB_DONE) return fail(LF_ERR_STATE, "job %llu has not finished", (unsigned long long)job);
    const size_t m = 3 * (size_t)j.n * sizeof(float);
    if (pos_xyz) std::memcpy(pos_xyz, j.rpos.data(), m);
    if (colour_rgb) std::memcpy(colour_rgb, j.rcol.data(), m);
    if (vel_xyz) std::memcpy(vel_xyz, j.rvel.data(), m);
    if (acc_xyz) std::memcpy(acc_xyz, j.racc.data(), m);
    if (latency_ms) *latency_ms = ms_since(j.t_submit, j.t_done);
    s->unreported.erase(std::remove(s->unreported.begin(), s->unreported.end(), job), s->unreported.end());
    s->jobs.erase(it);
    return LF_OK;
}

int lf_server_get_stats(const lf_server *s, lf_server_stats *out) {
    if (!s || !out) return fail(LF_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->admitted = s->admitted;
    out->finished = s->finished;
    out->queued = s->queue.size();
    out->running = s->running.size() + s->retiring.size();
    out->ticks = s->ticks;
    out->step_ticks = s->step_ticks;
    out->particle_steps = s->particle_steps;
    out->resident_particles = s->resident;
    out->slot_cap = s->pool.slot_cap();
    out->busy_seconds = s->busy_ms / 1e3;
    out->step_p50_ms = quantile(s->step_ms, 0.50);
    out->step_p99_ms = quantile(s->step_ms, 0.99);
    out->job_p50_ms = quantile(s->job_ms, 0.50);
    out->job_p99_ms = quantile(s->job_ms, 0.99);
    out->wait_p50_ms = quantile(s->wait_ms, 0.50);
    out->wait_p99_ms = quantile(s->wait_ms, 0.99);
    out->mean_occupancy =
        s->step_ticks ? s->occupancy_sum / (double)s->step_ticks / (double)s->pool.slot_cap() : 0.0;
    return LF_OK;
}

}  // extern "C"
