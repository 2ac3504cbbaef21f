// kamino_engine.hip -- batching and serving engine (kam_engine_*).
//
// One engine = one device, one stream, one HBM arena.  Active simulations of
// any supported shape share every step: the device batch holds their
// descriptors and tile lists, and a step is the four batched launches of
// DeviceBatch::launch, issued directly or replayed from a hipGraph captured per
// step parity.  With graph_steps = G > 1, runs of G steps during which no
// simulation finishes are captured as one graph of 4G launches, so the cost of
// a graph launch boundary is paid once per G steps; the steps between such runs
// go out as direct launches.  Membership changes
// (admission, retirement, cancellation) rebuild the tables and drop the graphs.  All device work is ordered on the engine stream, so a
// retired simulation's arena block can be handed to the next admission at
// once: the new upload is queued after the old download.
//
// With groups = G > 1 the active simulations are split into G groups of about
// equal cell count, each with its own batch, stream and graphs.  A step queues
// every group's launches on its own stream and then makes the engine stream
// wait for all of them, so the groups' kernels overlap on the device (one
// group's launch tails and ramps filled by the others' work) while uploads,
// result copies, the watchdog snapshot and step timing stay on the engine
// stream, ordered after every group's step.  After a membership change the
// group streams wait for the engine stream's uploads before their next launch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <unordered_map>
#include <vector>

#include "kamino_runtime.h"

namespace kam {

// First-fit allocator over one device allocation.
class Arena {
public:
    int init(size_t bytes) {
        bytes = bytes & ~(size_t)255;
        KAM_HIP(hipMalloc(&base_, bytes));
        size_ = bytes;
        free_[0] = bytes;
        return KAM_OK;
    }
    ~Arena() {
        if (base_) (void)hipFree(base_);
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes, o = it->first + bytes;
                free_.erase(it);
                if (rest) free_[o] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* base() const { return base_; }
    size_t size() const { return size_; }
    size_t used() const { return used_; }

private:
    char* base_ = nullptr;
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

// Pinned host buffers recycled by size.
class PinnedPool {
public:
    ~PinnedPool() {
        for (auto& kv : free_)
            for (void* p : kv.second) (void)hipHostFree(p);
    }
    void* get(size_t bytes) {
        auto& v = free_[bytes];
        if (!v.empty()) {
            void* p = v.back();
            v.pop_back();
            return p;
        }
        void* p = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
        return p;
    }
    void put(void* p, size_t bytes) { free_[bytes].push_back(p); }

private:
    std::map<size_t, std::vector<void*>> free_;
};

struct Request {
    uint64_t id = 0;
    kam_sim_desc sd{};
    Plan plan;
    Tables tables;
    SimDesc desc{};
    size_t off = 0, bytes = 0;
    int64_t total = 0, done = 0;
    int slot = -1;            // index of its watchdog word
    int group = 0;            // the group whose batch holds it
    int64_t admitStep = 0;    // engine step count at admission
    int state = KAM_REQ_ACTIVE;
    bool copying = false;  // retired; final copy to the host still in flight
    double admit_s = 0, done_s = 0;
    void* host = nullptr;  // pinned copy of the final fields
    size_t hostBytes = 0;
    hipEvent_t copied = nullptr;
};

}  // namespace kam

using namespace kam;

constexpr int kMaxGroups = 4;

struct kam_engine {
    kam_engine_desc ed{};
    hipStream_t stream = nullptr;  // the engine stream (group 0's stream)
    Arena arena;
    PinnedPool pinned;
    struct Group {
        DeviceBatch batch;
        hipStream_t stream = nullptr;  // group 0: the engine stream
        // [0][parity]: one step; [1][parity]: graph_steps steps
        hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
        hipEvent_t done = nullptr;     // groups > 0: recorded after each step's launches
        int64_t cells = 0;
    };
    Group groups[kMaxGroups];
    int ngroups = 1;
    hipEvent_t built = nullptr;     // engine stream, after a rebuild's uploads or a profiled step (groups > 1)
    std::unordered_map<uint64_t, Request> reqs;
    std::vector<uint64_t> active;     // batch order
    std::deque<uint64_t> draining;    // retired, final copy in flight (retirement order)
    std::vector<uint64_t> completed;  // final copy landed, not yet reported
    uint64_t nextId = 1;
    // watchdog: one word per slot on the device, snapshots on the host
    int* dStatus = nullptr;
    int* hStatus = nullptr;
    hipEvent_t statusEv = nullptr;
    int64_t statusStep = -1;          // engine step of the pending snapshot (-1: none)
    std::vector<int> freeSlots;
    int64_t failed = 0;
    int64_t totalSteps = 0;           // never reset
    int parity = 0;
    bool dirty = true;
    // replaced graphs wait here until their last launch has finished
    std::deque<std::pair<hipGraphExec_t, hipEvent_t>> oldGraphs;
    // statistics
    // Per step, not yet collected: the event recorded after it and, when the
    // GPU had gone idle before the step was queued, one recorded just before
    // it (else the step's interval starts at the previous step's end).
    struct StepEv {
        hipEvent_t start, end;
        int n;  // steps between the two events
    };
    std::deque<StepEv> stepEvents;
    std::vector<hipEvent_t> eventPool;
    hipEvent_t startEvent = nullptr;    // marks the beginning of the stats window
    bool windowOpen = false;
    std::vector<float> stepMs;
    int64_t steps = 0, cellSteps = 0, curCells = 0;
    int64_t admitted = 0, ncompleted = 0, rejected = 0, cancelled = 0;
    std::chrono::steady_clock::time_point t0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Retires the current graphs.  A graph exec owns its kernel arguments, so it
// is destroyed only once an event queued after its last launch has completed.
int drop_graphs(kam_engine* e) {
    for (int gi = 0; gi < e->ngroups; ++gi)
        for (auto& gp : e->groups[gi].graph)
            for (auto& g : gp) {
            if (!g) continue;
            hipEvent_t ev = nullptr;
            KAM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            KAM_HIP(hipEventRecord(ev, e->stream));
            e->oldGraphs.emplace_back(g, ev);
            g = nullptr;
            }
    return KAM_OK;
}

void reap_graphs(kam_engine* e, bool wait) {
    while (!e->oldGraphs.empty()) {
        auto& og = e->oldGraphs.front();
        if (wait) (void)hipEventSynchronize(og.second);
        else if (hipEventQuery(og.second) != hipSuccess) break;
        (void)hipGraphExecDestroy(og.first);
        (void)hipEventDestroy(og.second);
        e->oldGraphs.pop_front();
    }
}

int rebuild(kam_engine* e) {
    std::vector<SimDesc> sims[kMaxGroups];
    std::vector<Plan> plans[kMaxGroups];
    // tallest simulations first: their tiles start first
    std::stable_sort(e->active.begin(), e->active.end(), [e](uint64_t a, uint64_t b) {
        return e->reqs[a].plan.nT > e->reqs[b].plan.nT;
    });
    e->curCells = 0;
    for (int g = 0; g < e->ngroups; ++g) e->groups[g].cells = 0;
    for (uint64_t id : e->active) {
        Request& r = e->reqs[id];
        // pick up tuning-table changes; tile shapes never change values
        Plan fresh;
        if (make_plan(r.plan.nT, r.plan.nP, fresh) == KAM_OK) {
            r.plan = fresh;
            r.desc.advRows = fresh.advRows;
            r.desc.fftRows = fresh.fftRows;
            set_tile_cols(r.desc, fresh.advCols);
            int rc = attach_transform(r.desc, fresh, r.sd.precision, e->ed.device);
            if (rc != KAM_OK) return rc;
        }
        // the group with the fewest cells so far (tallest first: balanced)
        int g = 0;
        for (int k = 1; k < e->ngroups; ++k)
            if (e->groups[k].cells < e->groups[g].cells) g = k;
        r.group = g;
        sims[g].push_back(r.desc);
        plans[g].push_back(r.plan);
        const int64_t cells = (int64_t)r.plan.nT * r.plan.nP;
        e->groups[g].cells += cells;
        e->curCells += cells;
    }
    int rc = drop_graphs(e);
    if (rc != KAM_OK) return rc;
    // uploads go out on the engine stream, after the last step's join
    for (int g = 0; g < e->ngroups; ++g) {
        rc = e->groups[g].batch.build(sims[g], plans[g], e->stream);
        if (rc != KAM_OK) return rc;
    }
    if (e->ngroups > 1) {
        KAM_HIP(hipEventRecord(e->built, e->stream));
        for (int g = 1; g < e->ngroups; ++g) KAM_HIP(hipStreamWaitEvent(e->groups[g].stream, e->built, 0));
    }
    e->dirty = false;
    return KAM_OK;
}

// Step events time the steps and pace the host; no host read of device
// memory waits on them (result copies and the watchdog snapshot have events of
// their own), so they skip the system-scope release fence.
hipEvent_t take_event(kam_engine* e) {
    if (!e->eventPool.empty()) {
        hipEvent_t ev = e->eventPool.back();
        e->eventPool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return ev;
}

// Turns completed step events into per-step durations.  With `wait`, blocks
// on the stream first so every recorded step is collected.
int collect(kam_engine* e, bool wait) {
    if (wait) KAM_HIP(hipStreamSynchronize(e->stream));
    while (!e->stepEvents.empty()) {
        const kam_engine::StepEv se = e->stepEvents.front();
        hipEvent_t ev = se.end;
        if (!wait && hipEventQuery(ev) != hipSuccess) break;
        if (se.start || e->windowOpen) {
            float ms = 0.f;
            KAM_HIP(hipEventElapsedTime(&ms, se.start ? se.start : e->startEvent, ev));
            // a multi-step graph is timed as a whole: each of its steps gets the mean
            for (int k = 0; k < se.n; ++k) e->stepMs.push_back(ms / se.n);
        }
        if (se.start) e->eventPool.push_back(se.start);
        // the event becomes the start of the next interval
        if (e->startEvent) e->eventPool.push_back(e->startEvent);
        e->startEvent = ev;
        e->windowOpen = true;
        e->stepEvents.pop_front();
    }
    return KAM_OK;
}

int retire(kam_engine* e, Request& r, int state) {
    if (state == KAM_REQ_DONE) {
        const size_t N = (size_t)r.plan.nT * r.plan.nP;
        r.hostBytes = 3 * N * sizeof(float);
        r.host = e->pinned.get(r.hostBytes);
        if (!r.host) return set_err(KAM_ENOMEM, "pinned host allocation failed");
        float* h = static_cast<float*>(r.host);
        for (int f = 0; f < 3; ++f)
            KAM_HIP(hipMemcpyAsync(h + f * N, r.desc.state[e->parity][f], N * sizeof(float),
                                   hipMemcpyDeviceToHost, e->stream));
        KAM_HIP(hipEventCreateWithFlags(&r.copied, hipEventDisableTiming));
        KAM_HIP(hipEventRecord(r.copied, e->stream));
        r.copying = true;
        e->draining.push_back(r.id);
    }
    r.state = state;
    if (r.slot >= 0) e->freeSlots.push_back(r.slot);
    r.slot = -1;
    e->arena.release(r.off, r.bytes);
    release_tables(e->ed.device, r.sd.n_theta, r.sd.dt, r.sd.radius);
    e->active.erase(std::find(e->active.begin(), e->active.end(), r.id));
    e->dirty = true;
    return KAM_OK;
}

// Marks retired requests whose final copy has landed; stamps their time.
void drain(kam_engine* e) {
    while (!e->draining.empty()) {
        auto it = e->reqs.find(e->draining.front());
        if (it != e->reqs.end()) {
            Request& r = it->second;
            if (hipEventQuery(r.copied) != hipSuccess) break;
            r.copying = false;
            r.done_s = e->now();
            e->completed.push_back(r.id);
        }
        e->draining.pop_front();
    }
}

// Every 8 steps (at the first step boundary at or past each multiple of 8):
// quarantine requests whose watchdog word was set in the last snapshot, then
// take a new snapshot (copied on the stream, read 8 or more steps later).
int watchdog(kam_engine* e, int64_t before) {
    if (before / 8 == e->totalSteps / 8) return KAM_OK;
    if (e->statusStep >= 0 && hipEventQuery(e->statusEv) == hipSuccess) {
        std::vector<uint64_t> bad;
        for (uint64_t id : e->active) {
            const Request& r = e->reqs[id];
            if (r.admitStep < e->statusStep && (e->hStatus[r.slot] & 1)) bad.push_back(id);
        }
        for (uint64_t id : bad) {
            int rc = retire(e, e->reqs[id], KAM_REQ_FAILED);
            if (rc != KAM_OK) return rc;
            e->failed++;
        }
        e->statusStep = -1;
    }
    if (e->statusStep < 0 && !e->active.empty()) {
        KAM_HIP(hipMemcpyAsync(e->hStatus, e->dStatus, (size_t)e->ed.max_slots * sizeof(int),
                               hipMemcpyDeviceToHost, e->stream));
        KAM_HIP(hipEventRecord(e->statusEv, e->stream));
        e->statusStep = e->totalSteps;
    }
    return KAM_OK;
}

// Queues `n` steps of every active simulation: one step (n = 1), or one replay
// of the graph of graph_steps steps (n = graph_steps; the caller has checked
// that no active simulation finishes inside it).
int one_step(kam_engine* e, hipEvent_t* kernelEvents, int n = 1) {
    // keep at most max_inflight steps queued ahead of the GPU
    const size_t q = e->stepEvents.size(), lim = (size_t)e->ed.max_inflight;
    if (q >= lim) KAM_HIP(hipEventSynchronize(e->stepEvents[q - lim].end));
    if (e->dirty) {
        int rc = rebuild(e);
        if (rc != KAM_OK) return rc;
    }
    if (e->active.empty()) return KAM_OK;
    // an idle GPU: time this step from its own start, not from the last step's end
    hipEvent_t startEv = nullptr;
    if (e->stepEvents.empty() || hipEventQuery(e->stepEvents.back().end) == hipSuccess) {
        startEv = take_event(e);
        if (!startEv) return set_err(KAM_EHIP, "event creation failed");
        KAM_HIP(hipEventRecord(startEv, e->stream));
    }
    // every launch sequence flips the parity once per step
    const int np = (n & 1) ? e->parity ^ 1 : e->parity;
    const bool useGraph = e->ed.use_graph && !kernelEvents && (n > 1 || e->ed.graph_steps == 1);
    for (int gi = 0; gi < e->ngroups; ++gi) {
        kam_engine::Group& G = e->groups[gi];
        if (G.batch.nsims() == 0) continue;
        // profiling: every group on the engine stream, each with its own events
        hipStream_t st = kernelEvents ? e->stream : G.stream;
        if (useGraph) {
            hipGraphExec_t& g = G.graph[n > 1][e->parity];
            if (!g) {
                hipGraph_t graph = nullptr;
                KAM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                int rc = KAM_OK, p = e->parity;
                for (int k = 0; k < n && rc == KAM_OK; ++k) rc = G.batch.launch(KAM_PHASE_ALL, p, st, &p);
                hipError_t ce = hipStreamEndCapture(st, &graph);
                if (rc != KAM_OK) return rc;
                if (ce != hipSuccess) return set_err(KAM_EHIP, "graph capture: %s", hipGetErrorString(ce));
                ce = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ce != hipSuccess) return set_err(KAM_EHIP, "graph instantiate: %s", hipGetErrorString(ce));
            }
            KAM_HIP(hipGraphLaunch(g, st));
        } else {
            int p = e->parity;
            for (int k = 0; k < n; ++k) {
                int rc = G.batch.launch(KAM_PHASE_ALL, p, st, &p,
                                        kernelEvents ? kernelEvents + 2 * KC_COUNT * gi : nullptr);
                if (rc != KAM_OK) return rc;
            }
        }
        if (st != e->stream) {
            KAM_HIP(hipEventRecord(G.done, st));
            KAM_HIP(hipStreamWaitEvent(e->stream, G.done, 0));
        }
    }
    if (kernelEvents && e->ngroups > 1) {
        // the groups' states advanced on the engine stream: their own streams
        // wait for it before their next step
        KAM_HIP(hipEventRecord(e->built, e->stream));
        for (int g = 1; g < e->ngroups; ++g) KAM_HIP(hipStreamWaitEvent(e->groups[g].stream, e->built, 0));
    }
    e->parity = np;
    hipEvent_t ev = take_event(e);
    if (!ev) return set_err(KAM_EHIP, "event creation failed");
    KAM_HIP(hipEventRecord(ev, e->stream));
    e->stepEvents.push_back({startEv, ev, n});
    e->steps += n;
    e->cellSteps += n * e->curCells;
    // advance every request; retire the finished ones
    std::vector<uint64_t> finished;
    for (uint64_t id : e->active) {
        Request& r = e->reqs[id];
        r.done += n;
        if (r.done >= r.total) finished.push_back(id);
    }
    for (uint64_t id : finished) {
        int rc = retire(e, e->reqs[id], KAM_REQ_DONE);
        if (rc != KAM_OK) return rc;
        e->ncompleted++;
    }
    drain(e);
    reap_graphs(e, false);
    const int64_t before = e->totalSteps;
    e->totalSteps += n;
    int rc = watchdog(e, before);
    if (rc != KAM_OK) return rc;
    if (e->stepEvents.size() > 1024) return collect(e, false);
    return KAM_OK;
}

}  // namespace

extern "C" {

int kam_engine_create(const kam_engine_desc* d, kam_engine** out) {
    if (!d || !out) return set_err(KAM_EINVAL, "null argument");
    *out = nullptr;
    int rc = ensure_device(d->device);
    if (rc != KAM_OK) return rc;
    kam_engine* e = new kam_engine();
    e->ed = *d;
    if (e->ed.max_slots <= 0) e->ed.max_slots = 4096;
    if (e->ed.max_inflight <= 0) e->ed.max_inflight = 2;
    if (e->ed.graph_steps <= 0) e->ed.graph_steps = 1;
    if (e->ed.graph_steps > 64) {
        delete e;
        return set_err(KAM_EINVAL, "graph_steps %d above 64", d->graph_steps);
    }
    if (e->ed.groups <= 0) e->ed.groups = 1;
    if (e->ed.groups > kMaxGroups) {
        delete e;
        return set_err(KAM_EINVAL, "groups %d above %d", d->groups, kMaxGroups);
    }
    e->ngroups = e->ed.groups;
    e->t0 = std::chrono::steady_clock::now();
    size_t budget = d->hbm_budget_bytes;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete e;
            return set_err(KAM_EHIP, "hipMemGetInfo failed");
        }
        budget = fr / 2;
    }
    if ((rc = e->arena.init(budget)) != KAM_OK ||
        hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return rc != KAM_OK ? rc : set_err(KAM_EHIP, "stream creation failed");
    }
    e->groups[0].stream = e->stream;
    for (int g = 1; g < e->ngroups; ++g)
        if (hipStreamCreateWithFlags(&e->groups[g].stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e->groups[g].done, hipEventDisableTiming) != hipSuccess) {
            kam_engine_destroy(e);
            return set_err(KAM_EHIP, "group stream creation failed");
        }
    if (e->ngroups > 1 && hipEventCreateWithFlags(&e->built, hipEventDisableTiming) != hipSuccess) {
        kam_engine_destroy(e);
        return set_err(KAM_EHIP, "event creation failed");
    }
    const size_t sb = (size_t)e->ed.max_slots * sizeof(int);
    if (hipMalloc(&e->dStatus, sb) != hipSuccess || hipHostMalloc(&e->hStatus, sb, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&e->statusEv, hipEventDisableTiming) != hipSuccess ||
        hipMemset(e->dStatus, 0, sb) != hipSuccess) {
        kam_engine_destroy(e);
        return set_err(KAM_EHIP, "watchdog table allocation failed");
    }
    for (int k = e->ed.max_slots - 1; k >= 0; --k) e->freeSlots.push_back(k);
    *out = e;
    return KAM_OK;
}

int kam_engine_destroy(kam_engine* e) {
    if (!e) return KAM_OK;
    (void)hipSetDevice(e->ed.device);
    for (int g = 1; g < e->ngroups; ++g)
        if (e->groups[g].stream) (void)hipStreamSynchronize(e->groups[g].stream);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    reap_graphs(e, true);
    for (auto& kv : e->reqs) {
        Request& r = kv.second;
        if (r.state == KAM_REQ_ACTIVE) release_tables(e->ed.device, r.sd.n_theta, r.sd.dt, r.sd.radius);
        if (r.host) e->pinned.put(r.host, r.hostBytes);
        if (r.copied) (void)hipEventDestroy(r.copied);
    }
    for (const auto& se : e->stepEvents) {
        if (se.start) (void)hipEventDestroy(se.start);
        (void)hipEventDestroy(se.end);
    }
    for (hipEvent_t ev : e->eventPool) (void)hipEventDestroy(ev);
    if (e->startEvent) (void)hipEventDestroy(e->startEvent);
    if (e->statusEv) (void)hipEventDestroy(e->statusEv);
    if (e->dStatus) (void)hipFree(e->dStatus);
    if (e->hStatus) (void)hipHostFree(e->hStatus);
    for (int g = 1; g < e->ngroups; ++g) {
        if (e->groups[g].done) (void)hipEventDestroy(e->groups[g].done);
        if (e->groups[g].stream) (void)hipStreamDestroy(e->groups[g].stream);
    }
    if (e->built) (void)hipEventDestroy(e->built);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return KAM_OK;
}

int kam_engine_admit(kam_engine* e, const kam_sim_desc* d, const float* u_phi,
                     const float* u_theta, const float* rho, int64_t steps, uint64_t* id) {
    if (!e || !d || !u_phi || !u_theta || !id) return set_err(KAM_EINVAL, "null argument");
    if (steps <= 0) return set_err(KAM_EINVAL, "steps must be > 0");
    if (d->device != e->ed.device) return set_err(KAM_EINVAL, "simulation device differs from the engine's");
    if (!(d->dt > 0.0) || !(d->radius > 0.0)) return set_err(KAM_EINVAL, "bad dt / radius");
    KAM_HIP(hipSetDevice(e->ed.device));
    Request r;
    int rc = make_plan(d->n_theta, d->n_phi, r.plan);
    if (rc != KAM_OK) return rc;
    rc = check_precision(d->precision, r.plan);
    if (rc != KAM_OK) return rc;
    if ((int)e->active.size() >= e->ed.max_slots) {
        e->rejected++;
        return set_err(KAM_EAGAIN, "all %d slots busy", e->ed.max_slots);
    }
    r.bytes = kam_sim_bytes(d->n_theta, d->n_phi);
    if (!e->arena.alloc(r.bytes, &r.off)) {
        e->rejected++;
        return set_err(KAM_EAGAIN, "HBM budget exhausted (%zu of %zu bytes in use)", e->arena.used(),
                       e->arena.size());
    }
    rc = acquire_tables(e->ed.device, d->n_theta, d->dt, d->radius, &r.tables);
    if (rc != KAM_OK) {
        e->arena.release(r.off, r.bytes);
        return rc;
    }
    r.id = e->nextId++;
    r.sd = *d;
    r.total = steps;
    r.admit_s = e->now();
    r.admitStep = e->totalSteps;
    r.slot = e->freeSlots.back();
    e->freeSlots.pop_back();
    float* mem = reinterpret_cast<float*>(e->arena.base() + r.off);
    fill_desc(r.desc, r.plan, r.tables, *d, mem);
    rc = attach_transform(r.desc, r.plan, d->precision, e->ed.device);
    if (rc != KAM_OK) {
        e->freeSlots.push_back(r.slot);
        e->arena.release(r.off, r.bytes);
        release_tables(e->ed.device, d->n_theta, d->dt, d->radius);
        return rc;
    }
    r.desc.status = e->dStatus + r.slot;
    KAM_HIP(hipMemsetAsync(r.desc.status, 0, sizeof(int), e->stream));
    const size_t N = (size_t)d->n_theta * d->n_phi;
    // both state copies' u_theta padding row and rho default to zero
    KAM_HIP(hipMemsetAsync(mem, 0, 6 * N * sizeof(float), e->stream));
    KAM_HIP(hipMemcpyAsync(r.desc.state[e->parity][F_UPHI], u_phi, N * sizeof(float),
                           hipMemcpyHostToDevice, e->stream));
    KAM_HIP(hipMemcpyAsync(r.desc.state[e->parity][F_UTHETA], u_theta, (N - d->n_phi) * sizeof(float),
                           hipMemcpyHostToDevice, e->stream));
    if (rho)
        KAM_HIP(hipMemcpyAsync(r.desc.state[e->parity][F_RHO], rho, N * sizeof(float),
                               hipMemcpyHostToDevice, e->stream));
    *id = r.id;
    e->active.push_back(r.id);
    e->reqs[r.id] = r;
    e->admitted++;
    e->dirty = true;
    return KAM_OK;
}

int kam_engine_step(kam_engine* e, int nsteps) {
    if (!e || nsteps < 0) return set_err(KAM_EINVAL, "bad arguments");
    KAM_HIP(hipSetDevice(e->ed.device));
    const int G = e->ed.use_graph ? e->ed.graph_steps : 1;
    for (int k = 0; k < nsteps;) {
        // a run of G steps goes out as one graph when no simulation finishes inside it
        int n = 1;
        if (G > 1 && nsteps - k >= G && !e->active.empty()) {
            int64_t left = INT64_MAX;
            for (uint64_t id : e->active) {
                const Request& r = e->reqs[id];
                left = std::min(left, r.total - r.done);
            }
            if (left >= G) n = G;
        }
        int rc = one_step(e, nullptr, n);
        if (rc != KAM_OK) return rc;
        k += n;
    }
    return KAM_OK;
}

int kam_engine_profile(kam_engine* e, int nsteps, double* kernel_ms) {
    if (!e || nsteps <= 0) return set_err(KAM_EINVAL, "bad arguments");
    KAM_HIP(hipSetDevice(e->ed.device));
    // per group, an event before and after each kernel phase; the groups run
    // one after another on the engine stream and their times add up
    hipEvent_t kev[kMaxGroups * 2 * KC_COUNT];
    for (auto& ev : kev) KAM_HIP(hipEventCreate(&ev));
    double sum[KC_COUNT] = {0, 0, 0, 0};
    int rc = KAM_OK;
    for (int k = 0; k < nsteps && rc == KAM_OK; ++k) {
        rc = one_step(e, kev);
        if (rc != KAM_OK) break;
        KAM_HIP(hipStreamSynchronize(e->stream));
        for (int g = 0; g < e->ngroups; ++g) {
            if (e->groups[g].batch.nsims() == 0) continue;
            const hipEvent_t* ev = kev + 2 * KC_COUNT * g;
            for (int c = 0; c < KC_COUNT; ++c) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, ev[2 * c], ev[2 * c + 1]) == hipSuccess) sum[c] += ms;
            }
        }
    }
    for (auto& ev : kev) (void)hipEventDestroy(ev);
    if (rc == KAM_OK && kernel_ms)
        for (int c = 0; c < KC_COUNT; ++c) kernel_ms[c] = sum[c] / nsteps;
    return rc;
}

int kam_engine_poll(kam_engine* e, uint64_t id, kam_req_status* st) {
    if (!e || !st) return set_err(KAM_EINVAL, "null argument");
    auto it = e->reqs.find(id);
    if (it == e->reqs.end()) return set_err(KAM_ENOTFOUND, "unknown request %llu", (unsigned long long)id);
    drain(e);
    const Request& r = it->second;
    st->state = (r.state == KAM_REQ_DONE && r.copying) ? KAM_REQ_ACTIVE : r.state;
    st->reserved = 0;
    st->steps_done = r.done;
    st->steps_total = r.total;
    st->admit_s = r.admit_s;
    st->done_s = r.done_s;
    return KAM_OK;
}

int kam_engine_fetch(kam_engine* e, uint64_t id, float* u_phi, float* u_theta, float* rho) {
    if (!e) return set_err(KAM_EINVAL, "null engine");
    auto it = e->reqs.find(id);
    if (it == e->reqs.end()) return set_err(KAM_ENOTFOUND, "unknown request %llu", (unsigned long long)id);
    Request& r = it->second;
    if (r.state == KAM_REQ_FAILED) {
        e->reqs.erase(it);
        return set_err(KAM_ESTATE, "request %llu was quarantined (non-finite state)",
                       (unsigned long long)id);
    }
    if (r.state != KAM_REQ_DONE) return set_err(KAM_ESTATE, "request %llu has not finished",
                                                 (unsigned long long)id);
    KAM_HIP(hipEventSynchronize(r.copied));
    const size_t N = (size_t)r.plan.nT * r.plan.nP;
    const float* h = static_cast<const float*>(r.host);
    if (u_phi) std::memcpy(u_phi, h, N * sizeof(float));
    if (u_theta) std::memcpy(u_theta, h + N, (N - r.plan.nP) * sizeof(float));
    if (rho) std::memcpy(rho, h + 2 * N, N * sizeof(float));
    e->pinned.put(r.host, r.hostBytes);
    (void)hipEventDestroy(r.copied);
    e->reqs.erase(it);
    return KAM_OK;
}

int kam_engine_cancel(kam_engine* e, uint64_t id) {
    if (!e) return set_err(KAM_EINVAL, "null engine");
    auto it = e->reqs.find(id);
    if (it == e->reqs.end()) return set_err(KAM_ENOTFOUND, "unknown request %llu", (unsigned long long)id);
    if (it->second.state != KAM_REQ_ACTIVE) {
        // finished (its final copy may still be in flight), failed or
        // cancelled: give back its pinned copy and event, and forget it in the
        // completion list so that no later query reports it
        Request& r = it->second;
        if (r.copied) {
            KAM_HIP(hipEventSynchronize(r.copied));
            (void)hipEventDestroy(r.copied);
            r.copied = nullptr;
        }
        if (r.host) e->pinned.put(r.host, r.hostBytes);
        r.host = nullptr;
        e->completed.erase(std::remove(e->completed.begin(), e->completed.end(), id), e->completed.end());
        e->draining.erase(std::remove(e->draining.begin(), e->draining.end(), id), e->draining.end());
        e->reqs.erase(it);
        return KAM_OK;
    }
    int rc = retire(e, it->second, KAM_REQ_CANCELLED);
    e->cancelled++;
    e->reqs.erase(id);
    return rc;
}

int kam_engine_sync(kam_engine* e) {
    if (!e) return set_err(KAM_EINVAL, "null engine");
    KAM_HIP(hipSetDevice(e->ed.device));
    int rc = collect(e, true);
    drain(e);
    return rc;
}

int kam_engine_completed(kam_engine* e, uint64_t* ids, int max, int* count) {
    if (!e || !count || (max > 0 && !ids)) return set_err(KAM_EINVAL, "null argument");
    drain(e);
    const int n = std::min<int>(max, (int)e->completed.size());
    for (int k = 0; k < n; ++k) ids[k] = e->completed[k];
    e->completed.erase(e->completed.begin(), e->completed.begin() + n);
    *count = n;
    return KAM_OK;
}

int kam_engine_stats(kam_engine* e, kam_stats* s) {
    if (!e || !s) return set_err(KAM_EINVAL, "null argument");
    int rc = kam_engine_sync(e);
    if (rc != KAM_OK) return rc;
    std::memset(s, 0, sizeof *s);
    s->active = (int)e->active.size();
    s->steps = e->steps;
    s->cell_steps = e->cellSteps;
    std::vector<float> v(e->stepMs);
    double sum = 0;
    for (float x : v) sum += x;
    s->gpu_ms = sum;
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            const double pos = p * (v.size() - 1);
            const size_t lo = (size_t)std::floor(pos), hi = std::min(lo + 1, v.size() - 1);
            return (double)v[lo] + (pos - lo) * ((double)v[hi] - v[lo]);
        };
        s->step_p50_ms = pct(0.50);
        s->step_p99_ms = pct(0.99);
        s->step_max_ms = v.back();
    }
    s->admitted = e->admitted;
    s->completed = e->ncompleted;
    s->rejected = e->rejected;
    s->cancelled = e->cancelled;
    s->failed = e->failed;
    s->bytes_in_use = e->arena.used();
    s->bytes_budget = e->arena.size();
    return KAM_OK;
}

int kam_engine_reset_stats(kam_engine* e) {
    if (!e) return set_err(KAM_EINVAL, "null engine");
    KAM_HIP(hipSetDevice(e->ed.device));
    int rc = collect(e, true);
    if (rc != KAM_OK) return rc;
    // the next window starts at an event recorded now
    if (!e->startEvent) {
        e->startEvent = take_event(e);
        if (!e->startEvent) return set_err(KAM_EHIP, "event creation failed");
    }
    KAM_HIP(hipEventRecord(e->startEvent, e->stream));
    e->windowOpen = true;
    e->stepMs.clear();
    e->steps = 0;
    e->cellSteps = 0;
    e->admitted = e->ncompleted = e->rejected = e->cancelled = e->failed = 0;
    return KAM_OK;
}

double kam_engine_now(kam_engine* e) { return e ? e->now() : 0.0; }

int kam_tune_set(int n_theta, int adv_rows, int adv_cols, int fft_rows) {
    if (n_theta < 4 || adv_rows < 0 || adv_cols < 0 || fft_rows < 0)
        return set_err(KAM_EINVAL, "bad tuning entry");
    return tune_set(n_theta, adv_rows, adv_cols, fft_rows);
}

int kam_tune_set_transform(int n_theta, int transform) {
    if (n_theta < 4 || (transform != KAM_TRANSFORM_FFT && transform != KAM_TRANSFORM_DFT &&
                        transform != KAM_TRANSFORM_FFT8))
        return set_err(KAM_EINVAL, "bad transform entry");
    return tune_set_transform(n_theta, transform);
}

int kam_tune_get_transform(int n_theta, int* transform) {
    if (!transform) return set_err(KAM_EINVAL, "null argument");
    Plan p;
    int rc = make_plan(n_theta, 2 * n_theta, p);
    if (rc != KAM_OK) return rc;
    *transform = p.transform;
    return KAM_OK;
}

int kam_tune_get(int n_theta, int* adv_rows, int* adv_cols, int* fft_rows) {
    Plan p;
    int rc = make_plan(n_theta, 2 * n_theta, p);
    if (rc != KAM_OK) return rc;
    if (adv_rows) *adv_rows = p.advRows;
    if (adv_cols) *adv_cols = p.advCols;
    if (fft_rows) *fft_rows = p.fftRows;
    return KAM_OK;
}

int kam_autotune(int device, int n_theta, int batch, int steps, const char* cache_path, double* kernel_ms) {
    if (n_theta < 4 || batch < 1 || steps < 1) return set_err(KAM_EINVAL, "bad autotune arguments");
    Plan base;
    int rc = make_plan(n_theta, 2 * n_theta, base);
    if (rc != KAM_OK) return rc;
    kam_engine_desc ed{};
    ed.device = device;
    ed.max_slots = batch;
    ed.hbm_budget_bytes = (size_t)batch * (kam_sim_bytes(n_theta, 2 * n_theta) + 4096) + (1 << 20);
    ed.use_graph = 0;
    kam_engine* e = nullptr;
    if ((rc = kam_engine_create(&ed, &e)) != KAM_OK) return rc;
    const int nP = 2 * n_theta;
    std::vector<float> up((size_t)n_theta * nP), ut((size_t)(n_theta - 1) * nP), rho((size_t)n_theta * nP);
    kam_init_fields(n_theta, nP, 1.0, std::min(16, std::max(1, n_theta / 4)), 7, 1, up.data(), ut.data(), rho.data());
    kam_sim_desc d{};
    d.n_theta = n_theta;
    d.n_phi = nP;
    d.radius = 1.0;
    d.dt = 0.5 / n_theta;
    d.geometric = 1;
    d.precision = KAM_PREC_FP32;
    d.device = device;
    uint64_t id;
    for (int b = 0; b < batch && rc == KAM_OK; ++b)
        rc = kam_engine_admit(e, &d, up.data(), ut.data(), rho.data(), (int64_t)1 << 40, &id);
    // candidates: advection tiles, then FFT tile heights, each timed on its own kernels
    const int rowsC[4] = {4, 8, 12, 16}, colsC[2] = {64, 128};
    const int fftMax = base.fftRows;
    const int fftC[3] = {fftMax, std::max(1, fftMax / 2), std::max(1, fftMax / 4)};
    double best[KC_COUNT] = {1e30, 1e30, 1e30, 1e30};
    int bestR = base.advRows, bestW = base.advCols, bestF = base.fftRows;
    double k[KC_COUNT];
    for (int ci = 0; ci < 2 && rc == KAM_OK; ++ci)
        for (int ri = 0; ri < 4 && rc == KAM_OK; ++ri) {
            if (colsC[ci] > nP) continue;
            tune_set(n_theta, rowsC[ri], colsC[ci], bestF);
            e->dirty = true;
            if ((rc = kam_engine_profile(e, 2, k)) != KAM_OK) break;
            if ((rc = kam_engine_profile(e, steps, k)) != KAM_OK) break;
            if (k[KC_ADVECT] < best[KC_ADVECT]) {
                best[KC_ADVECT] = k[KC_ADVECT];
                bestR = rowsC[ri];
                bestW = colsC[ci];
            }
        }
    for (int fi = 0; fi < 3 && rc == KAM_OK; ++fi) {
        if (fi > 0 && fftC[fi] == fftC[fi - 1]) continue;
        tune_set(n_theta, bestR, bestW, fftC[fi]);
        e->dirty = true;
        if ((rc = kam_engine_profile(e, 2, k)) != KAM_OK) break;
        if ((rc = kam_engine_profile(e, steps, k)) != KAM_OK) break;
        const double f = k[KC_DIVFFT] + k[KC_IFFT];
        if (f < best[KC_DIVFFT] + best[KC_IFFT]) {
            best[KC_DIVFFT] = k[KC_DIVFFT];
            best[KC_IFFT] = k[KC_IFFT];
            best[KC_TRI] = k[KC_TRI];
            bestF = fftC[fi];
        }
    }
    // the phi transform: the row FFT (just timed) against its eight-wave form
    // and the MFMA DFT
    int bestT = KAM_TRANSFORM_FFT;
    if (rc == KAM_OK) {
        tune_set(n_theta, bestR, bestW, bestF);
        tune_set_transform(n_theta, KAM_TRANSFORM_FFT8);
        e->dirty = true;
        if ((rc = kam_engine_profile(e, 2, k)) == KAM_OK && (rc = kam_engine_profile(e, steps, k)) == KAM_OK &&
            k[KC_DIVFFT] + k[KC_IFFT] < best[KC_DIVFFT] + best[KC_IFFT]) {
            best[KC_DIVFFT] = k[KC_DIVFFT];
            best[KC_IFFT] = k[KC_IFFT];
            best[KC_TRI] = k[KC_TRI];
            bestT = KAM_TRANSFORM_FFT8;
        }
    }
    if (rc == KAM_OK && n_theta <= 1024) {
        tune_set(n_theta, bestR, bestW, bestF);
        tune_set_transform(n_theta, KAM_TRANSFORM_DFT);
        e->dirty = true;
        if ((rc = kam_engine_profile(e, 2, k)) == KAM_OK && (rc = kam_engine_profile(e, steps, k)) == KAM_OK &&
            k[KC_DIVFFT] + k[KC_IFFT] < best[KC_DIVFFT] + best[KC_IFFT]) {
            best[KC_DIVFFT] = k[KC_DIVFFT];
            best[KC_IFFT] = k[KC_IFFT];
            best[KC_TRI] = k[KC_TRI];
            bestT = KAM_TRANSFORM_DFT;
        }
    }
    kam_engine_destroy(e);
    tune_set_transform(n_theta, bestT);
    if (rc != KAM_OK) return rc;
    tune_set(n_theta, bestR, bestW, bestF);
    if (kernel_ms)
        for (int c = 0; c < KC_COUNT; ++c) kernel_ms[c] = best[c];
    if (cache_path && *cache_path) {
        FILE* f = std::fopen(cache_path, "a");
        if (!f) return set_err(KAM_EINVAL, "cannot append to %s", cache_path);
        std::fprintf(f, "gfx950 %d %d %d %d %d %.5f %.5f %d\n", n_theta, batch, bestR, bestW, bestF,
                     best[KC_ADVECT], best[KC_DIVFFT] + best[KC_IFFT], bestT);
        std::fclose(f);
    }
    return KAM_OK;
}

int kam_tune_load(const char* cache_path) {
    if (!cache_path) return set_err(KAM_EINVAL, "null path");
    FILE* f = std::fopen(cache_path, "r");
    if (!f) return set_err(KAM_ENOTFOUND, "no tuning cache at %s", cache_path);
    char line[256], arch[32];
    int nT, b, R, W, F, T;
    double t1, t2;
    int n = 0;
    // lines of 8 fields (written before the transform choice) or 9
    while (std::fgets(line, sizeof line, f)) {
        const int got = std::sscanf(line, "%31s %d %d %d %d %d %lf %lf %d", arch, &nT, &b, &R, &W, &F, &t1, &t2, &T);
        if (got >= 8 && std::strcmp(arch, "gfx950") == 0 && nT >= 4) {
            tune_set(nT, R, W, F);  // later lines win
            tune_set_transform(nT, got == 9 ? T : KAM_TRANSFORM_FFT);
            ++n;
        }
    }
    std::fclose(f);
    return n > 0 ? KAM_OK : set_err(KAM_ENOTFOUND, "no gfx950 entries in %s", cache_path);
}

}  // extern "C"
