"""The batched-frames and serving workloads on the device, at their configured sizes.

Batched frames: all 64 animation frames of the caustics scene (16 each at 128^2,
256^2, 512^2, 768^2; 2^20 photons per 128^2 of film, capped at 2^24) traced in one
server run; four frames' films are bit-identical to single gp_trace renders of the
same photons, and a 128^2 frame matches the oracle on the same photons.

Serving: a seeded Poisson stream (20 jobs/s) of 12 jobs over the 8 resident scenes
(the plain, caustics and large-mesh scenes, a batched-frames frame, and random
triangle soups of 1e5, 5e5, 1e6 and 2e6 triangles), films from {128..1024}^2 and
log-uniform budgets 2^20..2^26, submitted on arrival under an HBM budget that holds
one 1024^2 film at a time: every job's film is bit-identical to the same job traced
alone, and at every step the live jobs' footprints (as the server counts them) fit the
budget, with jobs waiting for room at some point."""
import time

import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import scenes, workloads
from tests.parity import image_metrics

pytestmark = pytest.mark.gpu


def test_c3_all_frames_in_one_server_run(gplib, orc):
    sds = [scenes.frames_c3(f) for f in range(64)]
    gss = [gp.GpuScene(sd, gplib) for sd in sds]
    print(f"[c3] {len(sds)} frames, {sum(sd.n_photons for sd in sds)} photons", flush=True)
    srv = gp.Server(hbm_budget_bytes=16 << 30, max_live_jobs=64, lib=gplib)
    ids = [srv.submit(gs, sd.camera, sd.n_photons, seed_base=f << 32, max_depth=16, tag=f)
           for f, (sd, gs) in enumerate(zip(sds, gss))]
    srv.step()
    assert srv.pending() == (0, 0)
    check = (0, 21, 47, 63)
    films = {}
    segs = 0
    for f, (j, sd) in enumerate(zip(ids, sds)):
        r, x, h = srv.take(j, sd.camera.res_x, sd.camera.res_y, film=f in check)
        assert r.status == 0 and r.user_tag == f and r.stats.photons == sd.n_photons
        segs += r.stats.segments
        if f in check:
            films[f] = (x, h)
    srv.close()
    print(f"[c3] batched segments {segs}", flush=True)
    for f in check:
        sd = sds[f]
        x, h, e, st = gp.render_forward(gss[f], sd.camera, sd.n_photons, seed_base=f << 32, hero_c=4,
                                        max_depth=16)
        assert np.array_equal(films[f][0], x) and np.array_equal(films[f][1], h), f
    # a 128^2 frame against the oracle on the same photons
    sd = sds[5]
    cx, ch, ce, cst = orc.render(orc.scene(sd), sd.camera, sd.n_photons, seed_base=5 << 32, hero_c=4,
                                 max_bounce=16, threads=8)
    gx, gh, ge, gst = gp.render_forward(gss[5], sd.camera, sd.n_photons, seed_base=5 << 32, hero_c=4,
                                        max_depth=16)
    m = image_metrics(gx, cx, sd.n_photons, sd.n_photons, res=(sd.camera.res_x, sd.camera.res_y))
    print(f"[c3] frame 5 vs oracle: {m}", flush=True)
    assert m["rel_rmse_max"] < 0.01 and abs(m["mean_lum_ratio"] - 1) < 0.003
    assert abs(gst["segments"] / cst["segments"] - 1) < 2e-3


def test_c4_poisson_serving_alone_equals_mix(gplib):
    sc = workloads.c4_scenes(gplib, full=True)
    print("[c4] scenes " + ", ".join(f"{n}:{sd.n_tris}" for n, sd, _ in sc), flush=True)
    # (each soup scene adds its light's two triangles)
    assert [sd.n_tris - 2 for n, sd, _ in sc[4:]] == [100_000, 500_000, 1_000_000, 2_000_000]
    jobs = workloads.c4_trace(12, lam=20.0, seed=4)
    # one 1024^2 film (32 MiB) plus tables fits, two do not
    budget = (1024 * 1024 * 32) + (8 << 20)
    srv = gp.Server(hbm_budget_bytes=budget, max_live_jobs=16, lib=gplib)
    t0 = time.perf_counter()
    sub = []          # (job id, job index) in submission order
    done = {}
    nxt = 0
    waited = False
    while nxt < len(jobs) or len(done) < len(jobs):
        now = time.perf_counter() - t0
        while nxt < len(jobs) and jobs[nxt]["arrival"] <= now:
            j = jobs[nxt]
            name, sd, gs = sc[j["scene"]]
            jid = srv.submit(gs, sd.with_camera_res(j["res"]).camera, j["photons"], seed_base=j["seed"],
                             max_depth=sd.max_depth, rr_depth=sd.rr_depth, tag=nxt)
            sub.append((jid, nxt))
            nxt += 1
        if len(sub) == len(done):
            time.sleep(0.002)
            continue
        srv.step(max_iterations=8)
        for jid in srv.poll():
            k = dict(sub)[jid]
            res = jobs[k]["res"]
            done[jid] = srv.take(jid, res, res, film=True)
        q, live, live_bytes = srv.pending(with_bytes=True)
        unfinished = [k for jid, k in sub if jid not in done]
        assert live + q <= len(unfinished)          # (the rest are retiring)
        # the admitted footprints never exceed the budget; they hold at least the films
        assert live_bytes <= budget and (live == 0 or live_bytes >= 128 * 128 * 32), (live, live_bytes)
        waited |= q > 0
    srv.close()
    print(f"[c4] 12 jobs served in {time.perf_counter() - t0:.2f} s; jobs waited for room: {waited}", flush=True)
    assert waited
    for jid, k in sub:
        j = jobs[k]
        name, sd, gs = sc[j["scene"]]
        r, x, h = done[jid]
        assert r.status == 0 and r.stats.photons == j["photons"]
        ax, ah, ae, ast = gp.render_forward(gs, sd.with_camera_res(j["res"]).camera, j["photons"],
                                            seed_base=j["seed"], hero_c=4, max_depth=sd.max_depth,
                                            rr_depth=sd.rr_depth)
        assert np.array_equal(x, ax) and np.array_equal(h, ah), (k, name)
        assert r.stats.segments == ast["segments"] and r.energy.emitted == ae["emitted"]
