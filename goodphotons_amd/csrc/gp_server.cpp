// Serving: continuous batching of render jobs in one resident wavefront, and the
// queue/grid autotuner.
//
// A job is (scene, camera, photon budget). Jobs wait in a FIFO until a device job
// slot and their HBM footprint (film + per-job tables) are free; they are then
// installed into the running wavefront's job table between iterations, and the
// scheduling kernel starts feeding their photons into free path slots on the next
// iteration, alongside the paths of jobs already in flight. A path emitted in
// iteration e finishes by iteration e + maxDepth - 1, so a job whose last photon
// was emitted in iteration e is complete once e + maxDepth iterations have run;
// its film is then final. Every photon is seeded by (seed_base + i, salt) of its own
// job and the film accumulates integers, so a job's film and ledger are
// bit-identical whether it ran alone or in the mix.
//
// The host never stalls the device between chunks of iterations:
//   - job installs and order changes are small kernels on the engine stream whose
//     data travel as kernel arguments (no host buffer has to outlive a copy);
//   - films are allocated and cleared stream-ordered (hipMallocAsync + fill kernel);
//   - after each chunk the control block is copied into a pinned ring slot and an
//     event recorded; the host reads a snapshot only once its event completes and
//     keeps at most kInflight chunks ahead of the device;
//   - a finished job's ledger rows are copied back on the engine stream (before its
//     slot can be reused) and its film on a second stream that overlaps the next
//     chunks, into pinned buffers recycled across jobs.
// A server is meant to be driven from one thread.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>

#include "gp_runtime.hpp"

using namespace gp;

namespace {

constexpr int kInflight = 2;          // chunks of iterations queued ahead of the host

// Pinned host buffers recycled by size (film downloads are a few fixed sizes).
struct PinnedPool {
    std::multimap<size_t, void*> freeList;
    std::map<void*, size_t> sizeOf;
    void* get(size_t n) {
        auto it = freeList.find(n);
        if (it != freeList.end()) {
            void* p = it->second;
            freeList.erase(it);
            return p;
        }
        void* p = nullptr;
        if (hipHostMalloc(&p, n) != hipSuccess) return nullptr;
        sizeOf[p] = n;
        return p;
    }
    void put(void* p) {
        if (p) freeList.emplace(sizeOf[p], p);
    }
    ~PinnedPool() {
        for (auto& kv : sizeOf) hipHostFree(kv.first);
    }
};

enum class JobState { Waiting, Live, Retiring, Finished };

}  // namespace

struct gp_server {
    gp_server_config cfg{};
    struct Job {
        int64_t id = 0;
        gp_job_desc desc{};
        gp_film* film = nullptr;
        int slot = -1;
        int64_t footprint = 0;
        gp_status status = GP_OK;
        JobState state = JobState::Waiting;
        double submitMs = 0, startMs = 0, finishMs = 0;
        double energyScale = 1.0;
        gp_energy energy{};
        gp_trace_stats stats{};
        int64_t installSeq = 0;                  // first chunk that ran with the job installed
        unsigned long long remaining = 0;        // photons not yet emitted (last snapshot)
        unsigned long long* accHost = nullptr;   // pinned, kSegs * kAccRow words
        long long* filmHost = nullptr;           // pinned, res * 4 words
        hipEvent_t accEv = nullptr, filmEv = nullptr;
    };
    std::map<int64_t, Job> jobs;
    std::deque<int64_t> waiting;
    std::vector<int64_t> slotJob;   // job id per slot, -1 free
    std::vector<int64_t> order;     // live job ids in emission priority order
    int64_t nextId = 1;
    int64_t liveBytes = 0;
    int64_t finished = 0;
    bool engineReady = false;
    uint64_t epoch = 0;                 // engine epoch of this server's control block
    // control-block snapshots of the chunks in flight
    struct Chunk {
        Ctrl* snap = nullptr;
        hipEvent_t ev = nullptr;
        int64_t seq = 0;
    };
    int64_t chunkSeq = 0;           // sequence number of the next chunk
    std::vector<Chunk> ring;
    std::deque<int> inflight;       // ring indices, oldest first
    int ringNext = 0;
    PinnedPool pool;
    std::vector<hipEvent_t> eventPool;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    hipEvent_t event() {
        if (!eventPool.empty()) {
            hipEvent_t e = eventPool.back();
            eventPool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        hipEventCreateWithFlags(&e, hipEventDisableTiming);
        return e;
    }
    void recycle(hipEvent_t& e) {
        if (e) eventPool.push_back(e);
        e = nullptr;
    }
};

namespace {

int64_t jobFootprint(const gp_job_desc& j) {
    return (int64_t)j.camera.res_x * j.camera.res_y * 32 + (int64_t)sizeof(DevJob) + 4096;
}

// Emission priority of the live jobs. With GP_ORDER_SHORTEST_FIRST the scheduling
// kernel fills free path slots for the job with the fewest photons left to emit
// first, so a small job admitted behind large ones finishes without waiting for them
// (their photons go into the slots left over); equal counts keep admission order.
// With GP_ORDER_ADMISSION the order is the admission order. Returns whether the order
// changed.
bool reorder(gp_server* srv) {
    if (srv->cfg.order_policy != GP_ORDER_SHORTEST_FIRST) return false;
    std::vector<int64_t> o = srv->order;
    std::stable_sort(o.begin(), o.end(), [&](int64_t a, int64_t b) {
        return srv->jobs[a].remaining < srv->jobs[b].remaining;
    });
    if (o == srv->order) return false;
    srv->order.swap(o);
    return true;
}

gp_status uploadOrder(Engine& E, gp_server* srv, int clearSlot) {
    OrderUpdate u{};
    u.nOrder = (int32_t)srv->order.size();
    int maxSlot = 0;
    for (size_t k = 0; k < srv->order.size(); ++k) {
        u.order[k] = srv->jobs[srv->order[k]].slot;
        maxSlot = std::max(maxSlot, u.order[k] + 1);
    }
    u.nJobs = maxSlot;
    u.clearSlot = clearSlot;
    launchSetOrder(E.q, u, device().stream);
    GP_HIP(hipGetLastError());
    return GP_OK;
}

// Film whose memory comes from the stream-ordered allocator of the engine stream.
gp_status streamFilm(const gp_camera_desc& cam, int64_t maxPhotons, gp_film** out) {
    hipStream_t s = device().stream;
    std::unique_ptr<gp_film> f(new gp_film());
    f->rx = cam.res_x;
    f->ry = cam.res_y;
    f->maxPhotons = std::max<int64_t>(1, maxPhotons);
    const size_t words = (size_t)f->rx * f->ry * 4;
    hipError_t e = hipMallocAsync((void**)&f->d, words * sizeof(unsigned long long), s);
    if (e != hipSuccess) return fail(GP_ERR_OOM, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    launchFill64(f->d, words, 0ull, s);
    GP_HIP(hipGetLastError());
    *out = f.release();
    return GP_OK;
}

void releaseFilm(gp_film* f, hipStream_t s) {
    if (!f) return;
    if (f->d) hipFreeAsync(f->d, s);
    f->d = nullptr;
    delete f;
}

gp_status admit(Engine& E, gp_server* srv) {
    bool changed = false;
    hipStream_t s = device().stream;
    while (!srv->waiting.empty()) {
        const int64_t id = srv->waiting.front();
        gp_server::Job& J = srv->jobs[id];
        int slot = -1;
        for (int k = 0; k < (int)srv->slotJob.size(); ++k)
            if (srv->slotJob[(size_t)k] < 0) { slot = k; break; }
        if (slot < 0) break;
        if (srv->liveBytes + J.footprint > srv->cfg.hbm_budget_bytes) break;
        // a scene with mirrors needs the mirror-vertex queue and k_mirror in the
        // iteration; both are set up at the start of a step, so wait for the next one
        if (J.desc.scene->nMirrors > 0 && !(E.q.mv[0] && E.mirrorActive)) break;
        srv->waiting.pop_front();
        JobInstall a{};
        gp_status st = streamFilm(J.desc.camera, J.desc.params.n_photons, &J.film);
        if (st == GP_OK) st = fillJob(J.desc.scene, &J.desc.camera, &J.film, 1, J.desc.params, slot, a.job, &a.cam);
        if (st != GP_OK) {
            J.status = st;
            J.state = JobState::Finished;
            J.startMs = J.finishMs = srv->now();
            srv->finished++;
            releaseFilm(J.film, s);
            J.film = nullptr;
            continue;
        }
        J.energyScale = a.job.energyScale;
        a.next = 0;
        a.end = (unsigned long long)J.desc.params.n_photons;
        a.slot = slot;
        launchInstallJob(E.q, a, s);
        GP_HIP(hipGetLastError());
        J.slot = slot;
        J.state = JobState::Live;
        J.installSeq = srv->chunkSeq;
        J.remaining = (unsigned long long)J.desc.params.n_photons;
        J.startMs = srv->now();
        srv->slotJob[(size_t)slot] = id;
        srv->order.push_back(id);
        srv->liveBytes += J.footprint;
        changed = true;
    }
    if (changed) {
        reorder(srv);
        return uploadOrder(E, srv, -1);
    }
    return GP_OK;
}

// A finished job's results leave the device: ledger rows on the engine stream (so a
// later install into the slot cannot overtake the copy), film on the copy stream
// once the chunk that finished it has completed.
gp_status startRetire(Engine& E, gp_server* srv, gp_server::Job& J, hipEvent_t chunkEv) {
    hipStream_t s = device().stream, cs = device().copyStream;
    J.accHost = (unsigned long long*)srv->pool.get(sizeof(unsigned long long) * kSegs * kAccRow);
    const size_t filmBytes = (size_t)J.film->rx * J.film->ry * 4 * sizeof(long long);
    J.filmHost = (long long*)srv->pool.get(filmBytes);
    if (!J.accHost || !J.filmHost) return fail(GP_ERR_OOM, "pinned host buffer allocation failed");
    J.accEv = srv->event();
    J.filmEv = srv->event();
    GP_HIP(hipMemcpyAsync(J.accHost, E.dJobAcc + (size_t)J.slot * kSegs * kAccRow,
                          sizeof(unsigned long long) * kSegs * kAccRow, hipMemcpyDeviceToHost, s));
    GP_HIP(hipEventRecord(J.accEv, s));
    GP_HIP(hipStreamWaitEvent(cs, chunkEv, 0));
    GP_HIP(hipMemcpyAsync(J.filmHost, J.film->d, filmBytes, hipMemcpyDeviceToHost, cs));
    GP_HIP(hipEventRecord(J.filmEv, cs));
    J.state = JobState::Retiring;
    return GP_OK;
}

void finishRetire(gp_server* srv, gp_server::Job& J) {
    unsigned long long ev[5] = {}, cv[kAccCounters] = {};
    sumAccRows(J.accHost, ev, cv);
    J.energy.emitted = (double)ev[0] / J.energyScale;
    J.energy.absorbed = (double)ev[1] / J.energyScale;
    J.energy.sensor = (double)ev[2] / J.energyScale;
    J.energy.escaped = (double)ev[3] / J.energyScale;
    J.energy.residual = (double)ev[4] / J.energyScale;
    addCounters(cv, J.stats);
    srv->pool.put(J.accHost);
    J.accHost = nullptr;
    srv->recycle(J.accEv);
    J.state = JobState::Finished;
    J.finishMs = srv->now();
    J.stats.wall_ms = J.finishMs - J.startMs;
    srv->finished++;
}

// Finalise retiring jobs whose ledger copy has landed (blocking: wait for them).
gp_status pollRetiring(gp_server* srv, bool blocking) {
    for (auto& kv : srv->jobs) {
        gp_server::Job& J = kv.second;
        if (J.state != JobState::Retiring) continue;
        if (blocking) {
            GP_HIP(hipEventSynchronize(J.accEv));
        } else {
            hipError_t q = hipEventQuery(J.accEv);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return fail(GP_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
        }
        finishRetire(srv, J);
    }
    return GP_OK;
}

// Read the oldest chunk's snapshot (waiting for it) and retire the jobs it shows done.
gp_status processOldest(Engine& E, gp_server* srv) {
    const int ri = srv->inflight.front();
    gp_server::Chunk& c = srv->ring[(size_t)ri];
    GP_HIP(hipEventSynchronize(c.ev));
    srv->inflight.pop_front();
    const Ctrl& h = *c.snap;
    if (h.overflow) return fail(GP_ERR_CAPACITY, "wavefront queue segment overflow");
    bool changed = false;
    for (size_t k = 0; k < srv->order.size();) {
        const int64_t id = srv->order[k];
        gp_server::Job& J = srv->jobs[id];
        // a snapshot taken before the job was installed shows the slot's previous job
        const int last = c.seq >= J.installSeq ? h.lastEmitIter[J.slot] : -1;
        if (last >= 0 && (int64_t)h.iter >= (int64_t)last + J.desc.params.max_depth) {
            gp_status st = startRetire(E, srv, J, c.ev);
            if (st != GP_OK) return st;
            srv->slotJob[(size_t)J.slot] = -1;
            srv->liveBytes -= J.footprint;
            srv->order.erase(srv->order.begin() + (long)k);
            changed = true;
            continue;
        }
        if (c.seq >= J.installSeq)
            J.remaining = h.endPhoton[J.slot] > h.nextPhoton[J.slot] ? h.endPhoton[J.slot] - h.nextPhoton[J.slot] : 0;
        ++k;
    }
    changed = reorder(srv) || changed;
    if (changed) return uploadOrder(E, srv, -1);
    return GP_OK;
}

gp_status launchChunk(Engine& E, gp_server* srv, int K) {
    hipStream_t s = device().stream;
    gp_status st = E.runIterations(K);
    if (st != GP_OK) return st;
    const int ri = srv->ringNext;
    srv->ringNext = (srv->ringNext + 1) % (int)srv->ring.size();
    gp_server::Chunk& c = srv->ring[(size_t)ri];
    GP_HIP(hipMemcpyAsync(c.snap, E.dCtrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
    GP_HIP(hipEventRecord(c.ev, s));
    c.seq = srv->chunkSeq++;
    srv->inflight.push_back(ri);
    return GP_OK;
}

}  // namespace

extern "C" {

gp_status gp_server_create(const gp_server_config* cfg, gp_server** out) {
    if (!cfg || !out) return fail(GP_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->max_live_jobs < 1 || cfg->max_live_jobs > kMaxJobs) return fail(GP_ERR_INVALID, "max_live_jobs must be 1..256");
    if (cfg->hbm_budget_bytes <= 0) return fail(GP_ERR_INVALID, "hbm_budget_bytes must be positive");
    if (cfg->order_policy != GP_ORDER_SHORTEST_FIRST && cfg->order_policy != GP_ORDER_ADMISSION)
        return fail(GP_ERR_INVALID, "unknown order_policy");
    gp_status st = requireDevice();
    if (st != GP_OK) return st;
    std::unique_ptr<gp_server> s(new gp_server());
    s->cfg = *cfg;
    s->slotJob.assign((size_t)cfg->max_live_jobs, -1);
    s->ring.resize(kInflight + 1);
    for (auto& c : s->ring) {
        GP_HIP(hipHostMalloc((void**)&c.snap, sizeof(Ctrl)));
        GP_HIP(hipEventCreateWithFlags(&c.ev, hipEventDisableTiming));
    }
    *out = s.release();
    return GP_OK;
}

gp_status gp_server_submit(gp_server* srv, const gp_job_desc* job, int64_t* jobId) {
    if (!srv || !job || !job->scene) return fail(GP_ERR_INVALID, "null argument");
    const gp_trace_params& p = job->params;
    if (p.n_photons < 0 || p.n_photons >= (int64_t(1) << 32)) return fail(GP_ERR_INVALID, "n_photons must be in [0, 2^32)");
    if (p.hero_c < 1 || p.hero_c > kMaxHero) return fail(GP_ERR_INVALID, "hero_c must be 1..4");
    if (p.max_depth < 1 || p.max_depth > 250) return fail(GP_ERR_INVALID, "max_depth must be 1..250");
    if (p.rr_depth < 0) return fail(GP_ERR_INVALID, "negative rr_depth");
    CamBasis cb;
    if (!cameraBasis(job->camera, cb)) return fail(GP_ERR_INVALID, "degenerate camera");
    const int64_t fp = jobFootprint(*job);
    if (fp > srv->cfg.hbm_budget_bytes) return fail(GP_ERR_REJECTED, "job footprint exceeds the server's HBM budget");
    gp_server::Job J;
    J.id = srv->nextId++;
    J.desc = *job;
    J.footprint = fp;
    J.submitMs = srv->now();
    srv->jobs[J.id] = J;
    srv->waiting.push_back(J.id);
    if (jobId) *jobId = J.id;
    return GP_OK;
}

gp_status gp_server_step(gp_server* srv, int32_t maxIterations, int32_t* finishedOut) {
    if (!srv) return fail(GP_ERR_INVALID, "null server");
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    if (E.owner && E.owner != srv) return fail(GP_ERR_STATE, "another server has jobs in flight on this device");
    if (!E.tunSet) { E.tun = defaultTuning(); E.tunSet = true; }
    gp_status st;
    int maxStackNeed = 0;
    bool mirrors = false;
    int64_t maxTris = 0;
    for (auto& kv : srv->jobs) {
        if (kv.second.state == JobState::Finished) continue;   // its scene is no longer used
        maxStackNeed = std::max(maxStackNeed, kv.second.desc.scene->stackNeed);
        mirrors = mirrors || kv.second.desc.scene->nMirrors > 0;
        maxTris = std::max(maxTris, kv.second.desc.scene->nTris);
    }
    const int lds = E.tun.lds_stack;
    if (srv->engineReady && !srv->inflight.empty()) {
        // a change of engine shape would re-allocate under running chunks
        const int need = std::max(0, maxStackNeed - lds);
        const bool connSpill = maxStackNeed > std::max(lds, 24);
        if (need > E.spillDepth || (need > 0) != (E.q.useSpill != 0) || connSpill != (E.q.connUseSpill != 0) ||
            (mirrors && !E.q.mv[0]) || mirrors != E.mirrorActive || binningFor(E.tun, maxTris) != (E.q.binned != 0)) {
            while (!srv->inflight.empty())
                if ((st = processOldest(E, srv)) != GP_OK) return st;
            GP_HIP(hipStreamSynchronize(device().stream));
        }
    }
    const unsigned oldCap = E.capacity, oldSeg = E.segCap;
    if ((st = E.ensure((unsigned)E.tun.queue_capacity, lds, maxStackNeed, 1, mirrors, maxTris)) != GP_OK) return st;
    E.mirrorActive = mirrors;
    const bool fresh = !srv->engineReady || E.capacity != oldCap || E.segCap != oldSeg || srv->epoch != E.epoch;
    if (fresh) {
        if (!srv->order.empty() || !srv->inflight.empty())
            return fail(GP_ERR_STATE, "engine re-allocated under a running server");
        E.initCtrl();
        GP_HIP(hipMemcpy(E.dCtrl, E.hCtrl, sizeof(Ctrl), hipMemcpyHostToDevice));
        E.parity = 0;
        srv->engineReady = true;
        srv->epoch = ++E.epoch;
    }
    E.owner = srv;   // recomputed when the step ends
    const int K = chunkIterations(E.tun);
    int64_t ran = 0;
    while (true) {
        if ((st = pollRetiring(srv, false)) != GP_OK) return st;
        if ((st = admit(E, srv)) != GP_OK) return st;
        const bool live = !srv->order.empty();
        if (!live && srv->inflight.empty()) break;
        if (live && (int)srv->inflight.size() < kInflight) {
            if ((st = launchChunk(E, srv, K)) != GP_OK) return st;
            ran += K;
        } else {
            if ((st = processOldest(E, srv)) != GP_OK) return st;
        }
        // everything else that has already completed, without waiting
        while (!srv->inflight.empty()) {
            hipError_t q = hipEventQuery(srv->ring[(size_t)srv->inflight.front()].ev);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) return fail(GP_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
            if ((st = processOldest(E, srv)) != GP_OK) return st;
        }
        if (maxIterations > 0 && ran >= maxIterations) break;
    }
    if (maxIterations <= 0 && (st = pollRetiring(srv, true)) != GP_OK) return st;
    // the engine stays this server's while it holds live jobs or chunks in flight
    E.owner = (srv->order.empty() && srv->inflight.empty()) ? nullptr : srv;
    if (finishedOut) *finishedOut = (int32_t)srv->finished;
    return GP_OK;
}

gp_status gp_server_take_result(gp_server* srv, int64_t jobId, gp_job_result* res, double* xyz,
                                double* hits) {
    if (!srv) return fail(GP_ERR_INVALID, "null server");
    auto it = srv->jobs.find(jobId);
    if (it == srv->jobs.end()) return fail(GP_ERR_INVALID, "unknown job id");
    gp_server::Job& J = it->second;
    if (J.state == JobState::Retiring) {
        GP_HIP(hipEventSynchronize(J.accEv));
        finishRetire(srv, J);
    }
    if (J.state != JobState::Finished) return fail(GP_ERR_STATE, "job not finished");
    if (res) {
        std::memset(res, 0, sizeof *res);
        res->job_id = J.id;
        res->user_tag = J.desc.user_tag;
        res->status = J.status;
        res->res_x = J.desc.camera.res_x;
        res->res_y = J.desc.camera.res_y;
        res->energy = J.energy;
        res->stats = J.stats;
        res->submit_ms = J.submitMs;
        res->start_ms = J.startMs;
        res->finish_ms = J.finishMs;
    }
    if (J.film) {
        if (J.filmEv) GP_HIP(hipEventSynchronize(J.filmEv));
        if (J.filmHost && (xyz || hits)) {
            const size_t np = (size_t)J.film->rx * J.film->ry;
            const double inv = J.film->scale > 0 ? 1.0 / J.film->scale : 0.0;
            const long long* w = J.filmHost;
            for (size_t i = 0; i < np; ++i) {
                if (xyz) {
                    xyz[3 * i + 0] = (double)w[4 * i + 0] * inv;
                    xyz[3 * i + 1] = (double)w[4 * i + 1] * inv;
                    xyz[3 * i + 2] = (double)w[4 * i + 2] * inv;
                }
                if (hits) hits[i] = (double)w[4 * i + 3];
            }
        }
        srv->pool.put(J.filmHost);
        J.filmHost = nullptr;
        srv->recycle(J.filmEv);
        releaseFilm(J.film, device().copyStream);
        J.film = nullptr;
    }
    srv->jobs.erase(it);
    return GP_OK;
}

gp_status gp_server_poll(gp_server* srv, int64_t* ids, int32_t capacity, int32_t* nOut) {
    if (!srv || (capacity > 0 && !ids)) return fail(GP_ERR_INVALID, "null argument");
    gp_status st = pollRetiring(srv, false);
    if (st != GP_OK) return st;
    std::vector<std::pair<double, int64_t>> done;
    for (auto& kv : srv->jobs)
        if (kv.second.state == JobState::Finished) done.emplace_back(kv.second.finishMs, kv.first);
    std::sort(done.begin(), done.end());
    int32_t n = 0;
    for (auto& d : done) {
        if (n >= capacity) break;
        ids[n++] = d.second;
    }
    if (nOut) *nOut = n;
    return GP_OK;
}

gp_status gp_server_pending(gp_server* srv, int32_t* queued, int32_t* live, int64_t* liveBytes) {
    if (!srv) return fail(GP_ERR_INVALID, "null server");
    if (queued) *queued = (int32_t)srv->waiting.size();
    if (live) *live = (int32_t)srv->order.size();
    if (liveBytes) *liveBytes = srv->liveBytes;
    return GP_OK;
}

void gp_server_destroy(gp_server* srv) {
    if (!srv) return;
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    hipStreamSynchronize(device().stream);
    hipStreamSynchronize(device().copyStream);
    // Leave no job in the shared engine's emission order.
    if (srv->engineReady && E.dCtrl) {
        srv->order.clear();
        uploadOrder(E, srv, -1);
        hipStreamSynchronize(device().stream);
    }
    for (auto& kv : srv->jobs) {
        gp_server::Job& J = kv.second;
        if (J.film) releaseFilm(J.film, device().stream);
        if (J.accEv) hipEventDestroy(J.accEv);
        if (J.filmEv) hipEventDestroy(J.filmEv);
    }
    hipStreamSynchronize(device().stream);
    for (auto& c : srv->ring) {
        if (c.snap) hipHostFree(c.snap);
        if (c.ev) hipEventDestroy(c.ev);
    }
    for (hipEvent_t e : srv->eventPool) hipEventDestroy(e);
    if (E.owner == srv) E.owner = nullptr;
    delete srv;
}

// ------------------------------------------------------------------ autotune
gp_status gp_autotune(gp_scene* scene, const gp_camera_desc* cam, const gp_trace_params* params,
                      gp_tuning* best) {
    if (!scene || !cam || !params) return fail(GP_ERR_INVALID, "null argument");
    gp_tuning base;
    gp_status st = gp_get_tuning(&base);
    if (st != GP_OK) return st;
    const int caps[] = {1 << 22, 1 << 23, 1 << 24};
    const int waves[] = {32, 64};
    const int ldss[] = {8, 12, 16};
    double bestRate = -1.0;
    gp_tuning bestT = base;
    // a failed trial leaves the caller's tuning in place
    struct Restore {
        const gp_tuning& t;
        bool armed = true;
        ~Restore() { if (armed) gp_set_tuning(&t); }
    } restore{base};
    auto trial = [&](const gp_tuning& t) -> gp_status {
        gp_status s0 = gp_set_tuning(&t);
        if (s0 != GP_OK) return s0;
        gp_trace_params p = *params;
        p.n_photons = std::min<int64_t>(params->n_photons, std::max<int64_t>(1 << 21, 2 * (int64_t)t.queue_capacity));
        gp_film* film = nullptr;
        if ((s0 = gp_film_create(cam->res_x, cam->res_y, 2 * p.n_photons, &film)) != GP_OK) return s0;
        gp_energy e{};
        gp_trace_stats s{};
        s0 = gp_trace(scene, cam, film, &p, &e, &s);        // warm (allocations, graph)
        if (s0 == GP_OK) s0 = gp_trace(scene, cam, film, &p, &e, &s);
        gp_film_destroy(film);
        if (s0 != GP_OK) return s0;
        const double rate = (double)s.segments / std::max(1e-6, s.wall_ms);
        if (rate > bestRate) { bestRate = rate; bestT = t; }
        return GP_OK;
    };
    for (int cap : caps)
        for (int w : waves)
            for (int l : ldss) {
                gp_tuning t = base;
                t.queue_capacity = cap;
                t.grid_waves_per_cu = w;
                t.lds_stack = l;
                if ((st = trial(t)) != GP_OK) return st;
            }
    // then ray binning off and on at the best shape (the grid ran with the caller's setting)
    const int sortFrom = bestT.ray_sort;
    for (int rs = 0; rs <= 1; ++rs) {
        if (rs == sortFrom) continue;
        gp_tuning t = bestT;
        t.ray_sort = rs;
        if ((st = trial(t)) != GP_OK) return st;
    }
    if ((st = gp_set_tuning(&bestT)) != GP_OK) return st;
    restore.armed = false;
    if (best) *best = bestT;
    return GP_OK;
}

}  // extern "C"
