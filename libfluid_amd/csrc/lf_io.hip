// lf_io.hip -- the data formats either side of the stepping path: exact
// state save/load (checkpoints), device-side scene setup and asynchronous
// .geo export (host code, members of lf::Pool, plus the checkpoint file
// format).

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstring>
#include <deque>
#include <fstream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lf_internal.h"
#include "lf_io.h"
#include "lf_pool.h"

#define LF_HIP(expr)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return lf::fail(LF_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace lf {

// ---------------------------------------------------------------------------
// exact state

int Pool::save_state(uint32_t id, std::vector<float> &out) {
    const SceneRec *s = scene(id);
    if (!s) return fail(LF_ERR_INVALID, "save_state: no scene %u", id);
    const size_t n = s->n;
    float *d = nullptr;
    if (int rc = dalloc(&d, 12 * n)) return rc;
    launch_state_out(stream_, make_scene_consts(id), P0_, L0_, V0_, A0_, d);
    out.resize(12 * n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out.data(), d, 12 * n * sizeof(float), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    dev_free(d);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "save_state: %s", hipGetErrorString(e));
    return LF_OK;
}

int Pool::load_state(uint32_t id, const float *in, uint64_t steps) {
    SceneRec *s = id < scenes_.size() && scenes_[id].live ? &scenes_[id] : nullptr;
    if (!s || !in) return fail(LF_ERR_INVALID, "load_state: bad scene %u or null data", id);
    const size_t n = s->n;
    float *d = nullptr;
    if (int rc = dalloc(&d, 12 * n)) return rc;
    hipError_t e = hipMemcpyAsync(d, in, 12 * n * sizeof(float), hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) {
        launch_state_in(stream_, s->first, s->n, d, P0_, L0_, V0_, A0_);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    dev_free(d);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "load_state: %s", hipGetErrorString(e));
    s->identity = true;
    s->fields = false;
    s->steps = steps;
    return LF_OK;
}

// ---------------------------------------------------------------------------
// device-side setup

int Pool::setup_jitter(uint32_t id, uint32_t id0, uint32_t nx, uint32_t ny, uint32_t nz, float dx, const float o[3],
                       float jitter, uint64_t seed) {
    const SceneRec *s = scene(id);
    if (!s || !o) return fail(LF_ERR_INVALID, "setup_jitter: bad scene %u or null origin", id);
    const uint64_t count = (uint64_t)nx * ny * nz;
    if ((uint64_t)id0 + count > s->n)
        return fail(LF_ERR_INVALID, "setup_jitter: ids [%u, %llu) exceed the scene's %u particles", id0,
                    (unsigned long long)(id0 + count), s->n);
    if (int rc = to_id_order(id)) return rc;
    launch_setup_jitter(stream_, s->first, id0, nx, ny, nz, dx, o, jitter, seed, P0_, L0_);
    LF_HIP(hipGetLastError());
    LF_HIP(hipStreamSynchronize(stream_));
    return LF_OK;
}

int Pool::setup_block(uint32_t id, uint32_t id0, uint32_t n, const float mn[3], const float mx[3], uint32_t *written,
                      float *mass, float *psize) {
    const SceneRec *s = scene(id);
    if (!s || !mn || !mx || n == 0) return fail(LF_ERR_INVALID, "setup_block: bad scene %u or arguments", id);
    if ((uint64_t)id0 + n > s->n)
        return fail(LF_ERR_INVALID, "setup_block: ids [%u, %llu) exceed the scene's %u particles", id0,
                    (unsigned long long)id0 + n, s->n);
    std::vector<float> ax[3];
    block_axes(n, mn, mx, s->params.rest_density, ax, mass, psize);
    const uint64_t full = (uint64_t)ax[0].size() * ax[1].size() * ax[2].size();
    const uint32_t count = (uint32_t)std::min<uint64_t>(n, full);
    if (written) *written = count;
    if (count == 0) return LF_OK;
    if (int rc = to_id_order(id)) return rc;
    const size_t tot = ax[0].size() + ax[1].size() + ax[2].size();
    float *d = nullptr;
    if (int rc = dalloc(&d, tot)) return rc;
    std::vector<float> h;
    h.reserve(tot);
    for (int a = 0; a < 3; ++a) h.insert(h.end(), ax[a].begin(), ax[a].end());
    hipError_t e = hipMemcpyAsync(d, h.data(), tot * sizeof(float), hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) {
        launch_setup_axes(stream_, s->first, id0, count, (uint32_t)ax[1].size(), (uint32_t)ax[2].size(), d,
                          d + ax[0].size(), d + ax[0].size() + ax[1].size(), P0_, L0_);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    dev_free(d);
    if (e != hipSuccess) return fail(LF_ERR_HIP, "setup_block: %s", hipGetErrorString(e));
    return LF_OK;
}

// ---------------------------------------------------------------------------
// asynchronous .geo export

struct Pool::Exporter {
    struct Buf {
        float *dev = nullptr, *host = nullptr;
        uint32_t *ids = nullptr;
        size_t cap = 0;  // particles
        hipEvent_t snap = nullptr, copied = nullptr;
    };
    struct Task {
        std::string base;
        uint32_t cnt = 0, n = 0;
        float pscale = 0.0f;
        int buf = -1;
    };
    hipStream_t copy = nullptr;
    std::vector<Buf> bufs;
    std::vector<int> free_bufs;
    std::deque<Task> queue;
    size_t pending = 0;
    bool stop = false;
    int err = LF_OK;
    std::string msg;
    std::mutex m;
    std::condition_variable cv, done;
    std::thread worker;

    void run() {
        for (;;) {
            Task t;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return stop || !queue.empty(); });
                if (queue.empty()) return;
                t = queue.front();
                queue.pop_front();
            }
            Buf &b = bufs[t.buf];  // bufs only grows on the submitting thread while no task is queued
            int rc = LF_OK;
            std::string why;
            if (hipEventSynchronize(b.copied) != hipSuccess) {
                rc = LF_ERR_HIP;
                why = "export copy failed";
            } else {
                std::vector<float> pos(3 * (size_t)t.n), col(3 * (size_t)t.n), vel(3 * (size_t)t.n);
                for (size_t k = 0; k < t.n; ++k)
                    for (int a = 0; a < 3; ++a) {
                        pos[3 * k + a] = b.host[12 * k + a];
                        col[3 * k + a] = b.host[12 * k + 3 + a];
                        vel[3 * k + a] = b.host[12 * k + 6 + a];
                    }
                rc = lf_write_geo(t.base.c_str(), t.cnt, pos.data(), col.data(), vel.data(), t.n, t.pscale);
                if (rc) why = lf_last_error();
            }
            std::lock_guard<std::mutex> lk(m);
            if (rc && err == LF_OK) {
                err = rc;
                msg = why;
            }
            free_bufs.push_back(t.buf);
            --pending;
            done.notify_all();
        }
    }
};

void Pool::stop_exporter() {
    if (!exporter_) return;
    {
        std::lock_guard<std::mutex> lk(exporter_->m);
        exporter_->stop = true;
    }
    exporter_->cv.notify_all();
    if (exporter_->worker.joinable()) exporter_->worker.join();
    for (auto &b : exporter_->bufs) {
        dev_free(b.dev);
        dev_free(b.ids);
        if (b.host) (void)hipHostFree(b.host);
        if (b.snap) (void)hipEventDestroy(b.snap);
        if (b.copied) (void)hipEventDestroy(b.copied);
    }
    if (exporter_->copy) (void)hipStreamDestroy(exporter_->copy);
    delete exporter_;
    exporter_ = nullptr;
}

int Pool::export_geo_async(uint32_t id, const char *filebase, uint32_t cnt, int order) {
    const SceneRec *s = scene(id);
    if (!s || !filebase) return fail(LF_ERR_INVALID, "export_geo_async: bad scene %u or null file name", id);
    if (order != LF_ORDER_CELL && order != LF_ORDER_ID) return fail(LF_ERR_INVALID, "export_geo_async: bad order");
    if (!exporter_) {
        exporter_ = new (std::nothrow) Exporter();
        if (!exporter_) return fail(LF_ERR_ALLOC, "export_geo_async: out of host memory");
        hipError_t e = hipStreamCreateWithFlags(&exporter_->copy, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete exporter_;
            exporter_ = nullptr;
            return fail(LF_ERR_HIP, "export_geo_async: %s", hipGetErrorString(e));
        }
        exporter_->worker = std::thread([ex = exporter_] { ex->run(); });
    }
    Exporter &x = *exporter_;
    const size_t n = s->n;
    int bi = -1;
    {
        std::lock_guard<std::mutex> lk(x.m);
        for (size_t k = 0; k < x.free_bufs.size(); ++k)
            if (x.bufs[x.free_bufs[k]].cap >= n) {
                bi = x.free_bufs[k];
                x.free_bufs.erase(x.free_bufs.begin() + k);
                break;
            }
    }
    if (bi < 0) {
        // A new buffer pair; the vector of buffers may only grow while the
        // worker holds no reference into it.
        if (int rc = export_wait()) return rc;
        Exporter::Buf b;
        b.cap = n;
        int rc = dalloc(&b.dev, 12 * n);
        if (!rc) rc = dalloc(&b.ids, n);
        hipError_t e = hipSuccess;
        if (!rc) e = hipHostMalloc((void **)&b.host, 12 * n * sizeof(float), hipHostMallocDefault);
        if (!rc && e == hipSuccess) e = hipEventCreateWithFlags(&b.snap, hipEventDisableTiming);
        if (!rc && e == hipSuccess) e = hipEventCreateWithFlags(&b.copied, hipEventDisableTiming);
        if (rc || e != hipSuccess) {
            dev_free(b.dev);
            dev_free(b.ids);
            if (b.host) (void)hipHostFree(b.host);
            if (b.snap) (void)hipEventDestroy(b.snap);
            if (b.copied) (void)hipEventDestroy(b.copied);
            if (rc) return rc;
            return fail(LF_ERR_ALLOC, "export_geo_async: %s", hipGetErrorString(e));
        }
        std::lock_guard<std::mutex> lk(x.m);
        x.bufs.push_back(b);
        bi = (int)x.bufs.size() - 1;
    }
    Exporter::Buf &b = x.bufs[bi];
    // Snapshot on the stepping stream, copy on the second stream.
    if (int rc = enqueue_export(id, b.dev, b.ids, order == LF_ORDER_ID)) return rc;
    LF_HIP(hipEventRecord(b.snap, stream_));
    LF_HIP(hipStreamWaitEvent(x.copy, b.snap, 0));
    LF_HIP(hipMemcpyAsync(b.host, b.dev, 12 * n * sizeof(float), hipMemcpyDeviceToHost, x.copy));
    LF_HIP(hipEventRecord(b.copied, x.copy));
    Exporter::Task t;
    t.base = filebase;
    t.cnt = cnt;
    t.n = (uint32_t)n;
    t.pscale = s->params.particle_size;
    t.buf = bi;
    {
        std::lock_guard<std::mutex> lk(x.m);
        x.queue.push_back(t);
        ++x.pending;
    }
    x.cv.notify_one();
    return LF_OK;
}

int Pool::export_wait() {
    if (!exporter_) return LF_OK;
    Exporter &x = *exporter_;
    std::unique_lock<std::mutex> lk(x.m);
    x.done.wait(lk, [&] { return x.pending == 0; });
    const int rc = x.err;
    const std::string msg = x.msg;
    x.err = LF_OK;
    x.msg.clear();
    lk.unlock();
    if (rc) return fail(rc, "asynchronous export: %s", msg.c_str());
    return LF_OK;
}

// ---------------------------------------------------------------------------
// checkpoint file: header, 12 floats per particle in id order (position high
// and low words, velocity, acceleration), FNV-1a 64 of all preceding bytes.

namespace {

constexpr char kMagic[8] = {'L', 'F', 'C', 'K', 'P', 'T', '0', '1'};

struct CkptHeader {
    char magic[8];
    uint32_t abi_version;
    uint32_t num_particles;
    uint64_t steps;
    float dt;
    uint32_t reserved;
    lf_params params;
};

uint64_t fnv1a(const void *p, size_t n, uint64_t h) {
    const unsigned char *c = (const unsigned char *)p;
    for (size_t k = 0; k < n; ++k) {
        h ^= c[k];
        h *= 0x100000001B3ull;
    }
    return h;
}

}  // namespace

int write_checkpoint(const char *path, const lf_params &p, uint32_t n, uint64_t steps, float dt,
                     const std::vector<float> &state) {
    if (!path) return fail(LF_ERR_INVALID, "checkpoint: null path");
    CkptHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.abi_version = LF_ABI_VERSION;
    h.num_particles = n;
    h.steps = steps;
    h.dt = dt;
    h.params = p;
    uint64_t sum = fnv1a(&h, sizeof(h), 0xCBF29CE484222325ull);
    sum = fnv1a(state.data(), state.size() * sizeof(float), sum);
    std::ofstream f(path, std::ios::binary);
    if (!f) return fail(LF_ERR_INVALID, "checkpoint: cannot open %s", path);
    f.write((const char *)&h, sizeof(h));
    f.write((const char *)state.data(), (std::streamsize)(state.size() * sizeof(float)));
    f.write((const char *)&sum, sizeof(sum));
    f.close();
    if (!f) return fail(LF_ERR_INVALID, "checkpoint: failed writing %s", path);
    return LF_OK;
}

int read_checkpoint(const char *path, lf_params *p, uint32_t *n, uint64_t *steps, float *dt,
                    std::vector<float> &state) {
    if (!path) return fail(LF_ERR_INVALID, "checkpoint: null path");
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(LF_ERR_INVALID, "checkpoint: cannot open %s", path);
    CkptHeader h;
    f.read((char *)&h, sizeof(h));
    if (!f || std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0)
        return fail(LF_ERR_INVALID, "checkpoint: %s is not a checkpoint", path);
    if (h.abi_version != LF_ABI_VERSION)
        return fail(LF_ERR_INVALID, "checkpoint: %s has ABI version %u, this library %d", path, h.abi_version,
                    LF_ABI_VERSION);
    if (h.num_particles == 0 || h.num_particles >= kMaxParticles)
        return fail(LF_ERR_INVALID, "checkpoint: %s holds a bad particle count", path);
    state.resize(12 * (size_t)h.num_particles);
    f.read((char *)state.data(), (std::streamsize)(state.size() * sizeof(float)));
    uint64_t sum = 0;
    f.read((char *)&sum, sizeof(sum));
    if (!f) return fail(LF_ERR_INVALID, "checkpoint: %s is truncated", path);
    uint64_t want = fnv1a(&h, sizeof(h), 0xCBF29CE484222325ull);
    want = fnv1a(state.data(), state.size() * sizeof(float), want);
    if (want != sum) return fail(LF_ERR_INVALID, "checkpoint: %s fails its checksum", path);
    if (int rc = lf_params_check(&h.params)) return rc;
    *p = h.params;
    *n = h.num_particles;
    *steps = h.steps;
    *dt = h.dt;
    return LF_OK;
}

}  // namespace lf
