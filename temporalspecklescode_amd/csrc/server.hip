// server.hip -- device-resident request loop with continuous batching.
//
// The reference can hold one simulation per process: its tables live in process-global
// `__constant__` symbols (Simulation.cuh:168-173 and every Interface/*.cuh). Here every request
// owns its tracer (medium, sources, sampler, buffers, HIP stream), so several requests are
// resident at once. A worker thread:
//   * admits queued requests in arrival order while the HBM footprint of the admitted set stays
//     within the budget and fewer than max_inflight are active;
//   * keeps up to three batches in flight per active request (one per stage of its tracer's
//     pipeline), round-robin over requests, so work that arrives while others run is interleaved
//     with it on the device;
//   * on completion copies the field into a small device buffer owned by the request and frees
//     the tracer, so finished results stay on the device until fetched;
//   * keeps the media it has built (grid, cross sections, walk grids) in a cache keyed by their
//     content, preloadable with sstmc_server_preload, so requests on a medium already resident
//     share it instead of copying and building it again.
// A request's result depends only on its config: batches, random numbers and reductions are
// per request and in a fixed order, so it is bit-identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/sstmc.h"
#include "config_copy.h"
#include "tracer.h"

using namespace sstmc;
typedef std::chrono::steady_clock Clock;

extern "C" void sstmc_destroy(sstmc_handle* h);
namespace sstmc {
NextEventEstimation* handleTracer(sstmc_handle* h);
int configValidate(const sstmc_config* c);
int mediumBuild(const sstmc_config* c, std::shared_ptr<Medium>* out);
int createShared(const sstmc_config* c, std::shared_ptr<Medium> medium, sstmc_handle** out);
}

namespace {

// What identifies a medium: dimensions, edges, material grid and cross sections. Its hash keys
// the cache; a hit is confirmed by comparing the whole identity.
struct MediumId {
    uint32_t dims = 0, nmat = 0;
    std::vector<double> axes[3], sigs, siga;
    std::vector<uint8_t> grid;
    uint64_t hash = 0;

    explicit MediumId(const sstmc_config& c) {
        const sstmc_medium& m = c.medium;
        dims = c.dims;
        nmat = m.materialNum;
        const double* ax[3] = {m.xAxis, m.yAxis, dims == 3 ? m.zAxis : nullptr};
        const uint32_t sz[3] = {m.xAxisSize, m.yAxisSize, dims == 3 ? m.zAxisSize : 0u};
        size_t ncell = 1;
        uint64_t h = 0x9E3779B97F4A7C15ull ^ dims;
        auto mix = [&h](uint64_t v) {
            h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
            h *= 0xff51afd7ed558ccdull;
            h ^= h >> 33;
        };
        for (int a = 0; a < 3; ++a) {
            if (ax[a]) axes[a].assign(ax[a], ax[a] + sz[a]);
            for (double e : axes[a]) { uint64_t b; std::memcpy(&b, &e, 8); mix(b); }
            mix(sz[a]);
            if (sz[a]) ncell *= sz[a] - 1;
        }
        grid.resize(ncell);
        for (size_t i = 0; i < ncell; ++i) grid[i] = (uint8_t)m.materialGrid[i];
        for (size_t i = 0; i + 8 <= ncell; i += 8) { uint64_t b; std::memcpy(&b, grid.data() + i, 8); mix(b); }
        for (size_t i = ncell & ~(size_t)7; i < ncell; ++i) mix(grid[i]);
        sigs.assign(m.sigs, m.sigs + nmat);
        siga.assign(nmat, 0.0);
        if (m.siga) siga.assign(m.siga, m.siga + nmat);
        for (uint32_t i = 0; i < nmat; ++i) {
            uint64_t b;
            std::memcpy(&b, &sigs[i], 8); mix(b);
            std::memcpy(&b, &siga[i], 8); mix(b);
        }
        hash = h;
    }
    bool operator==(const MediumId& o) const {
        if (hash != o.hash || dims != o.dims || nmat != o.nmat || grid != o.grid || sigs != o.sigs || siga != o.siga)
            return false;
        for (int a = 0; a < 3; ++a)
            if (axes[a] != o.axes[a]) return false;
        return true;
    }
};

struct CachedMedium {
    std::unique_ptr<MediumId> id;
    std::shared_ptr<Medium> medium;
    uint64_t lastUse = 0;
};

enum State { kQueued, kActive, kDone, kFailed, kFetched };

struct Request {
    uint64_t id;
    std::unique_ptr<ConfigCopy> cfg;
    std::shared_ptr<Medium> medium;    // from the cache (then cfg holds no grid)
    std::unique_ptr<MediumId> mediumId;   // when the medium is not cached yet
    State state = kQueued;
    int error = SSTMC_NO_ERROR;
    uint64_t footprint = 0;            // bytes counted in the server's inUse while active
    sstmc_handle* handle = nullptr;
    NextEventEstimation* tracer = nullptr;
    uint32_t nextBatch = 0, batches = 0;
    uint32_t renderings = 1, rendering = 0;   // renderingsNum; the one in flight
    uint64_t seed = 0;
    std::deque<hipEvent_t> inflight;
    double* result = nullptr;          // device: renderingsNum fields, u(l, v, t, i) like sstmc_render
    size_t fieldDoubles = 0;           // doubles of one rendering's field
    std::vector<double> iter;          // one per rendering
    uint64_t paths = 0, contributions = 0;
    Clock::time_point submitted, done;
};

}  // namespace

struct sstmc_server {
    int device = 0;
    uint64_t budget = 0, inUse = 0;
    uint32_t maxInflight = 4;
    std::mutex mu;
    std::condition_variable cv;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    std::deque<uint64_t> queue;
    std::vector<uint64_t> active;
    uint64_t nextId = 1;
    bool stop = false;
    std::thread worker;
    std::vector<double> latMs;
    uint64_t donePaths = 0, doneContrib = 0;
    Clock::time_point firstAdmit, lastDone;
    bool anyAdmitted = false;
    std::vector<CachedMedium> media;   // guarded by mu
    uint64_t useClock = 0;
    static constexpr size_t kMaxMedia = 16;

    std::shared_ptr<Medium> findMedium(const MediumId& id);
    void keepMedium(std::unique_ptr<MediumId> id, std::shared_ptr<Medium> m);
    void run();
    void admit();
    int pump(Request& r);
    void renderingDone(Request& r);
    void fail(Request& r, int rc);
    void release(Request& r);
    void account(Request& r);
};

static uint64_t estimate(const sstmc_config& c) {
    uint32_t P = c.simulation.batchSize ? c.simulation.batchSize : 1024;
    bool fp64 = c.precision.terms == SSTMC_TERMS_FP64;
    bool frames = false;
    for (uint32_t m = 1; m < c.medium.materialNum; ++m) {
        if (c.sampler.D[m] != 0 || (c.sampler.Ux && c.sampler.Ux[m] != 0) || (c.sampler.Uy && c.sampler.Uy[m] != 0) ||
            (c.sampler.Uz && c.sampler.Uz[m] != 0))
            frames = c.sampler.tNum > 1;
    }
    const uint64_t cells = (uint64_t)(c.medium.xAxisSize - 1) * (c.medium.yAxisSize - 1) *
                           (c.dims == 3 ? (uint64_t)(c.medium.zAxisSize - 1) : 1ull);
    const uint32_t R = c.simulation.renderingsNum ? c.simulation.renderingsNum : 1;
    return NextEventEstimation::footprint(P, c.illumination.num, c.view.num, c.sampler.tNum, fp64, frames, cells,
                                          c.tracer.isCBS != 0,
                                          c.sampler.type == SSTMC_SAMPLER_TEMPORAL_CORRELATION,
                                          (uint64_t)P * (c.simulation.fullIteration ? c.simulation.fullIteration : 1)) +
           16ull * c.illumination.num * c.view.num * c.sampler.tNum * R;   // the result buffer
}

// Brings the request's share of inUse up to what its tracer holds now (the vertex pool is sized
// from the first batch's exact vertex count, and grows after an overflow; the split-K workspaces
// are allocated by the first contraction). Caller holds mu.
void sstmc_server::account(Request& r) {
    if (!r.tracer) return;
    const uint64_t now = r.tracer->deviceBytes() + r.fieldDoubles * r.renderings * sizeof(double);
    if (now > r.footprint) {
        inUse += now - r.footprint;
        r.footprint = now;
    }
}

// Frees everything the request holds on the device except its finished result: the events of
// batches in flight (after they complete), the tracer. Caller holds mu.
void sstmc_server::release(Request& r) {
    for (hipEvent_t e : r.inflight) {
        hipEventSynchronize(e);
        hipEventDestroy(e);
    }
    r.inflight.clear();
    if (r.handle) sstmc_destroy(r.handle);
    r.handle = nullptr;
    r.tracer = nullptr;
}

void sstmc_server::fail(Request& r, int rc) {
    const bool wasActive = r.state == kActive;
    release(r);
    r.medium.reset();
    if (r.result) hipFree(r.result);
    r.result = nullptr;
    if (wasActive) {
        inUse -= r.footprint;
        active.erase(std::find(active.begin(), active.end(), r.id));
    }
    r.state = kFailed;
    r.error = rc ? rc : SSTMC_MISC_ERROR;
    r.done = Clock::now();
    cv.notify_all();
}

// A cached medium equal to `id`, or null. Caller holds mu.
std::shared_ptr<Medium> sstmc_server::findMedium(const MediumId& id) {
    for (CachedMedium& cm : media)
        if (*cm.id == id) {
            cm.lastUse = ++useClock;
            return cm.medium;
        }
    return nullptr;
}

// Adds a medium to the cache; when full, drops the least recently used one that no request holds.
// Caller holds mu.
void sstmc_server::keepMedium(std::unique_ptr<MediumId> id, std::shared_ptr<Medium> m) {
    if (findMedium(*id)) return;
    if (media.size() >= kMaxMedia) {
        size_t victim = media.size();
        for (size_t i = 0; i < media.size(); ++i)
            if (media[i].medium.use_count() == 1 && (victim == media.size() || media[i].lastUse < media[victim].lastUse))
                victim = i;
        if (victim == media.size()) return;   // all in use: not cached
        media.erase(media.begin() + victim);
    }
    media.push_back(CachedMedium{std::move(id), std::move(m), ++useClock});
}

void sstmc_server::admit() {
    // caller holds mu
    while (!queue.empty() && active.size() < maxInflight) {
        Request& r = *reqs[queue.front()];
        if (inUse + r.footprint > budget && !active.empty()) break;   // wait for memory to free up
        queue.pop_front();
        int rc = SSTMC_NO_ERROR;
        if (!r.medium) {
            // a medium submitted after this request, or preloaded since, may be cached by now
            r.medium = findMedium(*r.mediumId);
            if (!r.medium) {
                rc = mediumBuild(&r.cfg->c, &r.medium);
                if (!rc) keepMedium(std::move(r.mediumId), r.medium);
            }
        }
        sstmc_handle* h = nullptr;
        if (!rc) rc = createShared(&r.cfg->c, r.medium, &h);
        if (rc) {
            fail(r, rc);
            continue;
        }
        r.handle = h;
        r.tracer = handleTracer(h);
        r.batches = r.tracer->batches();
        r.nextBatch = 0;
        r.rendering = 0;
        r.fieldDoubles = 2ull * r.tracer->L() * r.tracer->V() * r.tracer->T();
        r.iter.assign(r.renderings, 0.0);
        r.state = kActive;
        inUse += r.footprint;
        active.push_back(r.id);
        if (hipMalloc((void**)&r.result, r.fieldDoubles * r.renderings * sizeof(double)) != hipSuccess) {
            r.result = nullptr;
            fail(r, SSTMC_ALLOCATION_ERROR);
            continue;
        }
        r.tracer->updateSeed(r.seed);   // rendering 0: the config's seed (MatlabInterface.cu:1799)
        if ((rc = r.tracer->begin())) {
            fail(r, rc);
            continue;
        }
        account(r);
        if (!anyAdmitted) { firstAdmit = Clock::now(); anyAdmitted = true; }
    }
}

// Queue batches (at most three in flight: one per pipeline stage) and retire finished ones.
// Returns 1 when the current rendering's last batch has completed, -1 on an error (r.error set).
int sstmc_server::pump(Request& r) {
    while (!r.inflight.empty()) {
        const hipError_t q = hipEventQuery(r.inflight.front());
        if (q == hipErrorNotReady) break;
        hipEventDestroy(r.inflight.front());
        r.inflight.pop_front();
        if (q != hipSuccess) { r.error = SSTMC_KERNEL_ERROR; return -1; }
    }
    while (r.nextBatch < r.batches && r.inflight.size() < 3) {
        int rc = r.tracer->enqueueBatch(r.nextBatch++);
        if (rc) { r.error = rc; return -1; }
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { r.error = SSTMC_DEVICE_ERROR; return -1; }
        if (hipEventRecord(e, r.tracer->doneStream()) != hipSuccess) {
            hipEventDestroy(e);
            r.error = SSTMC_DEVICE_ERROR;
            return -1;
        }
        r.inflight.push_back(e);
    }
    return (r.nextBatch == r.batches && r.inflight.empty()) ? 1 : 0;
}

// The current rendering's batches are done: finish() re-runs any batch that overflowed the vertex
// pool and copies the field into the request's result; then either the next rendering starts at
// seed + i (MatlabInterface.cu:1799) or the request completes and its tracer is freed.
void sstmc_server::renderingDone(Request& r) {
    int rc = r.tracer->finish(r.result + r.fieldDoubles * r.rendering, true);
    if (rc) {
        fail(r, rc);
        return;
    }
    account(r);
    r.iter[r.rendering] = (double)r.tracer->paths() + (double)r.tracer->getZeroContributionPaths();
    r.paths += r.tracer->paths();
    r.contributions += r.tracer->paths() * (uint64_t)r.tracer->L() * r.tracer->V() * r.tracer->T();
    if (++r.rendering < r.renderings) {
        r.tracer->updateSeed(r.seed + r.rendering);
        if ((rc = r.tracer->begin())) {
            fail(r, rc);
            return;
        }
        r.nextBatch = 0;
        return;
    }
    release(r);
    r.medium.reset();
    inUse -= r.footprint;
    active.erase(std::find(active.begin(), active.end(), r.id));
    r.done = Clock::now();
    r.state = kDone;
    latMs.push_back(std::chrono::duration<double, std::milli>(r.done - r.submitted).count());
    donePaths += r.paths;
    doneContrib += r.contributions;
    lastDone = r.done;
    cv.notify_all();
}

void sstmc_server::run() {
    hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (!stop) {
        admit();
        if (active.empty()) {
            cv.wait_for(lk, std::chrono::milliseconds(2));
            continue;
        }
        std::vector<uint64_t> ids = active;
        for (uint64_t id : ids) {
            Request& r = *reqs[id];
            if (r.state != kActive) continue;
            int st = pump(r);
            if (st < 0) fail(r, r.error);
            else if (st == 1) renderingDone(r);
        }
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        lk.lock();
    }
    for (uint64_t id : std::vector<uint64_t>(active)) release(*reqs[id]);
}

extern "C" {

int sstmc_server_create(uint32_t device, uint64_t hbm_budget_bytes, uint32_t max_inflight, sstmc_server** out) {
    if (!out) return SSTMC_INVALID_CONFIG;
    *out = nullptr;
    if (hipSetDevice((int)device) != hipSuccess) return SSTMC_DEVICE_ERROR;
    sstmc_server* s = new sstmc_server();
    s->device = (int)device;
    if (hbm_budget_bytes == 0) {
        size_t freeB = 0, totalB = 0;
        hipMemGetInfo(&freeB, &totalB);
        hbm_budget_bytes = (uint64_t)(freeB * 0.8);
    }
    s->budget = hbm_budget_bytes;
    s->maxInflight = max_inflight ? max_inflight : 4;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return SSTMC_NO_ERROR;
}

int sstmc_submit(sstmc_server* s, const sstmc_config* cfg, uint64_t* request_id) {
    if (!s || !cfg || !request_id) return SSTMC_INVALID_CONFIG;
    int rc = configValidate(cfg);
    if (rc) return rc;
    std::unique_ptr<Request> r(new Request());
    std::unique_ptr<MediumId> mid(new MediumId(*cfg));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->medium = s->findMedium(*mid);
    }
    r->cfg.reset(new ConfigCopy(*cfg, !r->medium));
    if (!r->medium) r->mediumId = std::move(mid);
    r->footprint = estimate(r->cfg->c);
    if (r->footprint > s->budget) return SSTMC_BUDGET_EXCEEDED;
    r->renderings = cfg->simulation.renderingsNum ? cfg->simulation.renderingsNum : 1;
    r->seed = cfg->simulation.seed;
    r->submitted = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->nextId++;
    *request_id = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    s->cv.notify_all();
    return SSTMC_NO_ERROR;
}

int sstmc_server_preload(sstmc_server* s, const sstmc_config* cfg) {
    if (!s || !cfg) return SSTMC_INVALID_CONFIG;
    std::unique_ptr<MediumId> mid(new MediumId(*cfg));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->findMedium(*mid)) return SSTMC_NO_ERROR;
    }
    std::shared_ptr<Medium> m;
    int rc = mediumBuild(cfg, &m);
    if (rc) return rc;
    // the walk grids of every axis, so no request waits for one
    for (int a = 0; a < (int)cfg->dims; ++a) {
        MediumView mv = m->view();
        if ((rc = m->walkGrid(a, &mv))) return rc;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->keepMedium(std::move(mid), std::move(m));
    return SSTMC_NO_ERROR;
}

int sstmc_poll(sstmc_server* s, uint64_t id) {
    if (!s) return SSTMC_INVALID_CONFIG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return SSTMC_INVALID_CONFIG;
    const Request& r = *it->second;
    if (r.state == kDone || r.state == kFetched) return SSTMC_NO_ERROR;
    if (r.state == kFailed) return r.error ? r.error : SSTMC_MISC_ERROR;
    return SSTMC_NOT_READY;
}

int sstmc_wait(sstmc_server* s, uint64_t id) {
    if (!s) return SSTMC_INVALID_CONFIG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return SSTMC_INVALID_CONFIG;
    Request& r = *it->second;
    s->cv.wait(lk, [&] { return r.state == kDone || r.state == kFailed || r.state == kFetched; });
    return r.state == kFailed ? (r.error ? r.error : SSTMC_MISC_ERROR) : SSTMC_NO_ERROR;
}

int sstmc_fetch(sstmc_server* s, uint64_t id, double* u, double* iter) {
    if (!s || !u) return SSTMC_INVALID_CONFIG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return SSTMC_INVALID_CONFIG;
    Request& r = *it->second;
    if (r.state == kFailed) return r.error ? r.error : SSTMC_MISC_ERROR;
    if (r.state != kDone) return SSTMC_NOT_READY;
    hipSetDevice(s->device);
    if (hipMemcpy(u, r.result, r.fieldDoubles * r.renderings * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return SSTMC_DEVICE_ERROR;
    if (iter)
        for (uint32_t i = 0; i < r.renderings; ++i) iter[i] = r.iter[i];
    hipFree(r.result);
    r.result = nullptr;
    r.state = kFetched;
    r.cfg.reset();
    return SSTMC_NO_ERROR;
}

int sstmc_server_stats_get(sstmc_server* s, sstmc_server_stats* out) {
    if (!s || !out) return SSTMC_INVALID_CONFIG;
    std::lock_guard<std::mutex> lk(s->mu);
    sstmc_server_stats st{};
    st.completed = s->latMs.size();
    if (!s->latMs.empty()) {
        std::vector<double> v = s->latMs;
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) {
            double pos = q * (v.size() - 1);
            size_t i = (size_t)pos;
            double f = pos - i;
            return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
        };
        st.p50_ms = pct(0.50);
        st.p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        st.mean_ms = sum / v.size();
        double busy = std::chrono::duration<double>(s->lastDone - s->firstAdmit).count();
        if (busy > 0) {
            st.paths_per_s = s->donePaths / busy;
            st.contributions_per_s = s->doneContrib / busy;
        }
    }
    st.hbm_in_use = s->inUse;
    st.hbm_budget = s->budget;
    *out = st;
    return SSTMC_NO_ERROR;
}

void sstmc_server_destroy(sstmc_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    hipSetDevice(s->device);
    for (auto& kv : s->reqs)
        if (kv.second->result) hipFree(kv.second->result);
    delete s;
}

}  // extern "C"
