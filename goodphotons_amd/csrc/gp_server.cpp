// Serving: continuous batching of render jobs in one resident wavefront, and the
// queue/grid autotuner.
//
// A job is (scene, camera, photon budget). Jobs wait in a FIFO until a device job
// slot and their HBM footprint (film + per-job tables) are free; they are then
// written into the running wavefront's job table between iterations, and the
// scheduling kernel starts feeding their photons into free path slots on the next
// iteration, alongside the paths of jobs already in flight. A path emitted in
// iteration e finishes by iteration e + maxDepth - 1, so a job whose last photon
// was emitted in iteration e is complete once e + maxDepth iterations have run;
// its film is then final and downloaded. Every photon is seeded by (seed_base + i,
// salt) of its own job and the film accumulates integers, so a job's film and
// ledger are bit-identical whether it ran alone or in the mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>

#include "gp_runtime.hpp"

using namespace gp;

struct gp_server {
    gp_server_config cfg{};
    struct Job {
        int64_t id = 0;
        gp_job_desc desc{};
        gp_film* film = nullptr;
        int slot = -1;
        int64_t footprint = 0;
        gp_status status = GP_OK;
        bool finished = false;
        double submitMs = 0, startMs = 0, finishMs = 0;
        double energyScale = 1.0;
        gp_energy energy{};
        gp_trace_stats stats{};
    };
    std::map<int64_t, Job> jobs;
    std::deque<int64_t> waiting;
    std::vector<int64_t> slotJob;   // job id per slot, -1 free
    std::vector<int64_t> order;     // live job ids in admission order
    int64_t nextId = 1;
    int64_t liveBytes = 0;
    bool engineReady = false;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

int64_t jobFootprint(const gp_job_desc& j) {
    return (int64_t)j.camera.res_x * j.camera.res_y * 32 + (int64_t)sizeof(DevJob) + 4096;
}

// Writes the slot's job record and control fields (stream-ordered between iterations).
gp_status uploadSlot(Engine& E, int slot) {
    hipStream_t s = device().stream;
    Ctrl* h = E.hCtrl;
    Ctrl* d = E.dCtrl;
    GP_HIP(hipMemcpyAsync(E.dJobs + slot, E.hJobs + slot, sizeof(DevJob), hipMemcpyHostToDevice, s));
    GP_HIP(hipMemcpyAsync(&d->nextPhoton[slot], &h->nextPhoton[slot], sizeof(unsigned long long), hipMemcpyHostToDevice, s));
    GP_HIP(hipMemcpyAsync(&d->endPhoton[slot], &h->endPhoton[slot], sizeof(unsigned long long), hipMemcpyHostToDevice, s));
    GP_HIP(hipMemcpyAsync(&d->lastEmitIter[slot], &h->lastEmitIter[slot], sizeof(int32_t), hipMemcpyHostToDevice, s));
    GP_HIP(hipMemcpyAsync(&d->active[slot], &h->active[slot], sizeof(int32_t), hipMemcpyHostToDevice, s));
    return E.resetJobAcc(slot);
}

gp_status uploadOrder(Engine& E, gp_server* srv) {
    hipStream_t s = device().stream;
    Ctrl* h = E.hCtrl;
    h->nOrder = (int32_t)srv->order.size();
    for (size_t k = 0; k < srv->order.size(); ++k) h->order[k] = srv->jobs[srv->order[k]].slot;
    int maxSlot = 0;
    for (size_t k = 0; k < srv->order.size(); ++k) maxSlot = std::max(maxSlot, h->order[k] + 1);
    h->nJobs = maxSlot;
    GP_HIP(hipMemcpyAsync(E.dCtrl->order, h->order, sizeof(int32_t) * kMaxJobs, hipMemcpyHostToDevice, s));
    GP_HIP(hipMemcpyAsync(&E.dCtrl->nOrder, &h->nOrder, sizeof(int32_t), hipMemcpyHostToDevice, s));
    GP_HIP(hipMemcpyAsync(&E.dCtrl->nJobs, &h->nJobs, sizeof(int32_t), hipMemcpyHostToDevice, s));
    // The host copy of `order` must stay unchanged until this copy has run.
    GP_HIP(hipStreamSynchronize(s));
    return GP_OK;
}

gp_status admit(Engine& E, gp_server* srv) {
    bool changed = false;
    while (!srv->waiting.empty()) {
        const int64_t id = srv->waiting.front();
        gp_server::Job& J = srv->jobs[id];
        int slot = -1;
        for (int k = 0; k < (int)srv->slotJob.size(); ++k)
            if (srv->slotJob[(size_t)k] < 0) { slot = k; break; }
        if (slot < 0) break;
        if (srv->liveBytes + J.footprint > srv->cfg.hbm_budget_bytes) break;
        srv->waiting.pop_front();
        gp_status st = gp_film_create(J.desc.camera.res_x, J.desc.camera.res_y,
                                      std::max<int64_t>(1, J.desc.params.n_photons), &J.film);
        if (st == GP_OK) {
            DevJob dj;
            st = fillJob(J.desc.scene, J.desc.camera, J.film, J.desc.params, slot, dj);
            if (st == GP_OK) {
                E.hJobs[slot] = dj;
                J.energyScale = dj.energyScale;
            }
        }
        if (st != GP_OK) {
            J.status = st;
            J.finished = true;
            J.startMs = J.finishMs = srv->now();
            if (J.film) { gp_film_destroy(J.film); J.film = nullptr; }
            continue;
        }
        Ctrl* h = E.hCtrl;
        h->nextPhoton[slot] = 0;
        h->endPhoton[slot] = (unsigned long long)J.desc.params.n_photons;
        h->lastEmitIter[slot] = J.desc.params.n_photons == 0 ? (int32_t)h->iter : -1;
        h->active[slot] = 1;
        if ((st = uploadSlot(E, slot)) != GP_OK) return st;
        J.slot = slot;
        J.startMs = srv->now();
        srv->slotJob[(size_t)slot] = id;
        srv->order.push_back(id);
        srv->liveBytes += J.footprint;
        changed = true;
    }
    if (changed) return uploadOrder(E, srv);
    return GP_OK;
}

gp_status retire(Engine& E, gp_server* srv, int32_t* finished) {
    const Ctrl& h = *E.hCtrl;
    bool changed = false;
    for (size_t k = 0; k < srv->order.size();) {
        const int64_t id = srv->order[k];
        gp_server::Job& J = srv->jobs[id];
        const int slot = J.slot;
        const int last = h.lastEmitIter[slot];
        if (last >= 0 && (int64_t)h.iter >= (int64_t)last + J.desc.params.max_depth) {
            unsigned long long ev[5], cv[4];
            gp_status st = E.readJobAcc(slot, ev, cv);
            if (st != GP_OK) return st;
            J.energy.emitted = (double)ev[0] / J.energyScale;
            J.energy.absorbed = (double)ev[1] / J.energyScale;
            J.energy.sensor = (double)ev[2] / J.energyScale;
            J.energy.escaped = (double)ev[3] / J.energyScale;
            J.energy.residual = (double)ev[4] / J.energyScale;
            J.stats.photons = cv[0];
            J.stats.segments = cv[1];
            J.stats.connections = cv[2];
            J.stats.splats = cv[3];
            J.finished = true;
            J.finishMs = srv->now();
            J.stats.wall_ms = J.finishMs - J.startMs;
            E.hCtrl->active[slot] = 0;
            srv->slotJob[(size_t)slot] = -1;
            srv->liveBytes -= J.footprint;
            srv->order.erase(srv->order.begin() + (long)k);
            changed = true;
            continue;
        }
        ++k;
    }
    if (changed) {
        GP_HIP(hipMemcpy(E.dCtrl->active, E.hCtrl->active, sizeof(int32_t) * kMaxJobs, hipMemcpyHostToDevice));
        gp_status st = uploadOrder(E, srv);
        if (st != GP_OK) return st;
    }
    int32_t n = 0;
    for (auto& kv : srv->jobs) n += kv.second.finished ? 1 : 0;
    if (finished) *finished = n;
    return GP_OK;
}

}  // namespace

extern "C" {

gp_status gp_server_create(const gp_server_config* cfg, gp_server** out) {
    if (!cfg || !out) return fail(GP_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->max_live_jobs < 1 || cfg->max_live_jobs > kMaxJobs) return fail(GP_ERR_INVALID, "max_live_jobs must be 1..256");
    if (cfg->hbm_budget_bytes <= 0) return fail(GP_ERR_INVALID, "hbm_budget_bytes must be positive");
    gp_status st = requireDevice();
    if (st != GP_OK) return st;
    std::unique_ptr<gp_server> s(new gp_server());
    s->cfg = *cfg;
    s->slotJob.assign((size_t)cfg->max_live_jobs, -1);
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
    if (!E.tunSet) { E.tun = defaultTuning(); E.tunSet = true; }
    gp_status st;
    int maxDepthScene = 0;
    for (auto& kv : srv->jobs) maxDepthScene = std::max(maxDepthScene, kv.second.desc.scene->depth);
    const int lds = E.tun.lds_stack;
    const unsigned oldCap = E.capacity, oldSeg = E.segCap;
    if ((st = E.ensure((unsigned)E.tun.queue_capacity, lds, std::max(0, maxDepthScene + 1 - lds))) != GP_OK) return st;
    const bool fresh = !srv->engineReady || E.capacity != oldCap || E.segCap != oldSeg;
    if (fresh) {
        E.initCtrl();
        GP_HIP(hipMemcpy(E.dCtrl, E.hCtrl, sizeof(Ctrl), hipMemcpyHostToDevice));
        E.parity = 0;
        srv->engineReady = true;
    }
    const int K = std::max(2, E.tun.iters_per_graph & ~1);
    int64_t ran = 0;
    while (true) {
        if ((st = admit(E, srv)) != GP_OK) return st;
        if (srv->order.empty()) break;
        if ((st = E.runIterations(K)) != GP_OK) return st;
        ran += K;
        if ((st = E.readCtrl()) != GP_OK) return st;
        if ((st = retire(E, srv, finishedOut)) != GP_OK) return st;
        if (maxIterations > 0 && ran >= maxIterations) break;
    }
    int32_t n = 0;
    for (auto& kv : srv->jobs) n += kv.second.finished ? 1 : 0;
    if (finishedOut) *finishedOut = n;
    return GP_OK;
}

gp_status gp_server_take_result(gp_server* srv, int64_t jobId, gp_job_result* res, double* xyz,
                                double* hits) {
    if (!srv) return fail(GP_ERR_INVALID, "null server");
    auto it = srv->jobs.find(jobId);
    if (it == srv->jobs.end()) return fail(GP_ERR_INVALID, "unknown job id");
    gp_server::Job& J = it->second;
    if (!J.finished) return fail(GP_ERR_STATE, "job not finished");
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
    gp_status st = GP_OK;
    if (J.film) {
        if (xyz || hits) st = gp_film_download(J.film, xyz, hits);
        gp_film_destroy(J.film);
        J.film = nullptr;
    }
    srv->jobs.erase(it);
    return st;
}

gp_status gp_server_pending(gp_server* srv, int32_t* queued, int32_t* live) {
    if (!srv) return fail(GP_ERR_INVALID, "null server");
    if (queued) *queued = (int32_t)srv->waiting.size();
    if (live) *live = (int32_t)srv->order.size();
    return GP_OK;
}

void gp_server_destroy(gp_server* srv) {
    if (!srv) return;
    for (auto& kv : srv->jobs)
        if (kv.second.film) gp_film_destroy(kv.second.film);
    // Leave no job active in the shared engine.
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    if (E.hCtrl && E.dCtrl) {
        for (int k = 0; k < kMaxJobs; ++k) E.hCtrl->active[k] = 0;
        E.hCtrl->nOrder = 0;
        hipMemcpy(E.dCtrl->active, E.hCtrl->active, sizeof(int32_t) * kMaxJobs, hipMemcpyHostToDevice);
        hipMemcpy(&E.dCtrl->nOrder, &E.hCtrl->nOrder, sizeof(int32_t), hipMemcpyHostToDevice);
    }
    delete srv;
}

// ------------------------------------------------------------------ autotune
gp_status gp_autotune(gp_scene* scene, const gp_camera_desc* cam, const gp_trace_params* params,
                      gp_tuning* best) {
    if (!scene || !cam || !params) return fail(GP_ERR_INVALID, "null argument");
    gp_tuning base;
    gp_status st = gp_get_tuning(&base);
    if (st != GP_OK) return st;
    const int caps[] = {1 << 20, 1 << 21, 1 << 22};
    const int waves[] = {16, 32};
    const int ldss[] = {8, 16};
    double bestRate = -1.0;
    gp_tuning bestT = base;
    for (int cap : caps)
        for (int w : waves)
            for (int l : ldss) {
                gp_tuning t = base;
                t.queue_capacity = cap;
                t.grid_waves_per_cu = w;
                t.lds_stack = l;
                if ((st = gp_set_tuning(&t)) != GP_OK) return st;
                gp_trace_params p = *params;
                p.n_photons = std::min<int64_t>(params->n_photons, std::max<int64_t>(1 << 21, 2 * (int64_t)cap));
                gp_film* film = nullptr;
                if ((st = gp_film_create(cam->res_x, cam->res_y, 2 * p.n_photons, &film)) != GP_OK) return st;
                gp_energy e{};
                gp_trace_stats s{};
                st = gp_trace(scene, cam, film, &p, &e, &s);        // warm (allocations, graph)
                if (st == GP_OK) st = gp_trace(scene, cam, film, &p, &e, &s);
                gp_film_destroy(film);
                if (st != GP_OK) return st;
                const double rate = (double)s.segments / std::max(1e-6, s.wall_ms);
                if (rate > bestRate) { bestRate = rate; bestT = t; }
            }
    if ((st = gp_set_tuning(&bestT)) != GP_OK) return st;
    if (best) *best = bestT;
    return GP_OK;
}

}  // extern "C"
