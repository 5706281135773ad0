"""Workload drivers for the large-mesh, batched-frames and serving configurations.

  python -m goodphotons_amd.workloads c2 [--photons N] [--level L]
      icosphere (20*4^L tris) dielectric on a floor, 1024^2 film, depth 12.
  python -m goodphotons_amd.workloads c3
      64 animation frames of the caustics scene (16 each at 128^2..768^2, budget
      2^20*px/128^2 capped at 2^24) traced one at a time (gp_trace per frame) and
      all together in one wavefront (gp_server); reports segments/s of both.
  python -m goodphotons_amd.workloads c4 [--jobs J]
      Poisson arrivals (20 jobs/s, seeded) over 8 resident scenes (c0, c1, c2, a c3
      frame, 4 random soups of 1e5..2e6 tris), films from {128,256,512,1024}^2,
      log-uniform budgets 2^20..2^26: continuous-batching server (shortest remaining
      job first, and in admission order) vs one job at a time, both with the
      reference's plain-wavefront scheduling (2^20 slots, host read per pass, no graph)
      and with this engine's tuned gp_trace; under the arrival trace and saturated;
      p50/p99 latency and aggregate segments/s.
Each prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

import goodphotons_amd as gp
from goodphotons_amd import scenes


def _pct(a, q):
    return float(np.percentile(np.asarray(a, float), q)) if len(a) else None


# ---------------------------------------------------------------- c2
def run_c2(photons=1 << 28, level=9, res=1024, profile=True):
    lib = gp.load()
    t0 = time.perf_counter()
    sd = scenes.icosphere_c2(level=level, res=res)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    gs = gp.GpuScene(sd, lib)
    t_up = time.perf_counter() - t0
    info = gs.info()
    film = gp.Film(res, res, max_photons=photons, lib=lib)
    gp.trace(gs, sd.camera, film, 1 << 20, hero_c=4, max_depth=sd.max_depth)   # warm-up
    film.reset()
    t0 = time.perf_counter()
    e, st = gp.trace(gs, sd.camera, film, photons, hero_c=4, max_depth=sd.max_depth)
    dt = time.perf_counter() - t0
    st = st.as_dict()
    xyz, hits = film.download()
    out = {"workload": f"c2 icosphere level {level} ({info['n_tris']} tris), {res}^2, {photons} photons, depth 12",
           "segments_per_s": st["segments"] / dt, "seconds": dt, "segments": st["segments"],
           "connections": st["connections"], "photons": st["photons"], "iterations": st["iterations"],
           "scene_gen_s": t_gen, "scene_create_s": t_up, "bvh_build_ms": info["build_ms"],
           "bvh_depth": info["bvh_depth"], "device_bytes": info["device_bytes"],
           "lit_pixels": int((xyz[:, 1] > 0).sum()), "film_finite": bool(np.isfinite(xyz).all()),
           "ledger": e.as_dict()}
    if profile:
        t = lib.tuning()
        lib.set_tuning(profile=1)
        try:
            film.reset()
            _, pst = gp.trace(gs, sd.camera, film, min(photons, 1 << 24), hero_c=4, max_depth=sd.max_depth)
        finally:
            lib.set_tuning(profile=0)
        p = pst.as_dict()
        out["kernel_ms_profiled_2^24"] = dict(zip(gp.abi.KERNEL_NAMES, p["kernel_ms"]))
        out["tuning"] = {k: getattr(t, k) for k, _ in t._fields_}
    return out


# ---------------------------------------------------------------- c3
def run_c3(frames=64, autotune=True):
    lib = gp.load()
    sds = [scenes.frames_c3(f) for f in range(frames)]
    gss = [gp.GpuScene(sd, lib) for sd in sds]
    tuned = None
    if autotune:
        import ctypes as C
        t = gp.abi.Tuning()
        p = gp.trace_params(1 << 22, hero_c=4, max_depth=16)
        lib.check(lib.lib.gp_autotune(gss[40].h, C.byref(sds[40].camera), C.byref(p), C.byref(t)))
        tuned = t.as_dict()
    # warm-up (graph capture, allocations)
    gp.render_forward(gss[0], sds[0].camera, 1 << 20, hero_c=4, max_depth=16)
    # one at a time
    t0 = time.perf_counter()
    plain_segs = 0
    plain_films = {}
    for f, (sd, gs) in enumerate(zip(sds, gss)):
        x, h, e, st = gp.render_forward(gs, sd.camera, sd.n_photons, seed_base=f << 32, hero_c=4,
                                        max_depth=16)
        plain_segs += st["segments"]
        if f in (0, 21, 47, 63):
            plain_films[f] = x
    t_plain = time.perf_counter() - t0
    # all frames in one wavefront
    srv = gp.Server(hbm_budget_bytes=16 << 30, max_live_jobs=max(64, frames), lib=lib)
    t0 = time.perf_counter()
    ids = [srv.submit(gs, sd.camera, sd.n_photons, seed_base=f << 32, max_depth=16, tag=f)
           for f, (sd, gs) in enumerate(zip(sds, gss))]
    srv.step()
    t_batch_gpu = time.perf_counter() - t0
    batch_segs = 0
    identical = True
    for f, (j, sd) in enumerate(zip(ids, sds)):
        r, x, h = srv.take(j, sd.camera.res_x, sd.camera.res_y, film=f in plain_films)
        batch_segs += r.stats.segments
        if f in plain_films:
            identical &= bool(np.array_equal(x, plain_films[f]))
    t_batch = time.perf_counter() - t0
    srv.close()
    return {"workload": f"c3 {frames} animated caustics frames (128^2..768^2, 2^20..2^24 photons)",
            "plain_segments_per_s": plain_segs / t_plain, "plain_seconds": t_plain,
            "batched_segments_per_s": batch_segs / t_batch, "batched_seconds": t_batch,
            "batched_gpu_seconds": t_batch_gpu, "speedup": (batch_segs / t_batch) / (plain_segs / t_plain),
            "segments": batch_segs, "films_identical_plain_vs_batched": identical, "autotuned": tuned}


# ---------------------------------------------------------------- c4
def c4_scenes(lib, full=True):
    defs = [("c0", lambda r: scenes.cornell_c0(r)), ("c1", lambda r: scenes.caustics_c1(r)),
            ("c2", lambda r: scenes.icosphere_c2(level=9 if full else 6, res=r)),
            ("c3f7", lambda r: scenes.caustics_c1(r, frame=7))]
    sizes = [100_000, 500_000, 1_000_000, 2_000_000] if full else [20_000, 40_000, 60_000, 80_000]
    for k, n in enumerate(sizes):
        defs.append((f"soup{n}", (lambda n, k: (lambda r: scenes.soup(n, 100 + k, res=r)))(n, k)))
    out = []
    for name, fn in defs:
        sd = fn(128)
        out.append((name, sd, gp.GpuScene(sd, lib)))
    return out


def c4_trace(n_jobs, lam=20.0, seed=4):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / lam, n_jobs))
    jobs = []
    for i in range(n_jobs):
        jobs.append({"arrival": float(t[i]), "scene": int(rng.integers(0, 8)),
                     "res": int(rng.choice([128, 256, 512, 1024])),
                     "photons": int(round(2.0 ** rng.uniform(20, 26))), "seed": i << 32})
    return jobs


def _cam(sd, res):
    return sd.with_camera_res(res).camera


# The reference's plain wavefront (src/render_cuda.cu:13623-13679): one job at a
# time, a pool of min(N, 2^20) path slots, and a blocking host read of the live count
# after every pass; no graph. Reproduced with this engine's kernels by its tuning.
REFERENCE_STYLE_TUNING = {"queue_capacity": 1 << 20, "use_graph": 0, "iters_per_graph": 1}


def _summary(lat, segs, t_first, t_end):
    span = t_end - t_first
    return {"p50_latency_s": _pct(lat, 50), "p99_latency_s": _pct(lat, 99),
            "segments_per_s": segs / span, "makespan_s": span}


def _run_plain(lib, sc, jobs, arrivals, tuning=None):
    """One job at a time in arrival order: gp_trace, then its film downloaded."""
    saved = {k: getattr(lib.tuning(), k) for k in (tuning or {})}
    if tuning:
        lib.set_tuning(**tuning)
    try:
        t0 = time.perf_counter()
        lat, segs = [], 0
        for j, arr in zip(jobs, arrivals):
            now = time.perf_counter() - t0
            if now < arr:
                time.sleep(arr - now)
            name, sd, gs = sc[j["scene"]]
            x, h, e, st = gp.render_forward(gs, _cam(sd, j["res"]), j["photons"], seed_base=j["seed"],
                                            hero_c=4, max_depth=sd.max_depth, rr_depth=sd.rr_depth)
            segs += st["segments"]
            lat.append(time.perf_counter() - t0 - arr)
        return _summary(lat, segs, arrivals[0], time.perf_counter() - t0)
    finally:
        if tuning:
            lib.set_tuning(**saved)


def _run_server(lib, sc, jobs, arrivals, order_policy=gp.abi.ORDER_SHORTEST_FIRST, max_live_jobs=64):
    """Continuous batching: submit on arrival, keep the device fed, take each result
    (film downloaded) as soon as it finishes. max_live_jobs = 1 runs one job at a time
    through the same loop (no overlap of one job's drain with the next job's photons)."""
    srv = gp.Server(hbm_budget_bytes=8 << 30, max_live_jobs=max_live_jobs, lib=lib, order_policy=order_policy)
    try:
        t0 = time.perf_counter()
        live = {}
        lat, segs = [None] * len(jobs), 0
        nxt = 0
        while nxt < len(jobs) or live:
            now = time.perf_counter() - t0
            while nxt < len(jobs) and arrivals[nxt] <= now:
                j = jobs[nxt]
                name, sd, gs = sc[j["scene"]]
                jid = srv.submit(gs, _cam(sd, j["res"]), j["photons"], seed_base=j["seed"],
                                 max_depth=sd.max_depth, rr_depth=sd.rr_depth, tag=nxt)
                live[jid] = (nxt, j["res"])
                nxt += 1
            if not live:
                time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
                continue
            srv.step(max_iterations=8)
            for jid in srv.poll():
                idx, res = live.pop(jid)
                r, x, h = srv.take(jid, res, res, film=True)
                lat[idx] = time.perf_counter() - t0 - arrivals[idx]
                segs += r.stats.segments
        return _summary(lat, segs, arrivals[0], time.perf_counter() - t0)
    finally:
        srv.close()


def run_c4(n_jobs=60, lam=20.0, full=True):
    """The serve loop against two one-job-at-a-time baselines, under the Poisson trace
    (latency; throughput there is bounded by the arrival rate when the device keeps
    up) and saturated (every job queued at t=0: throughput capacity)."""
    lib = gp.load()
    sc = c4_scenes(lib, full)
    jobs = c4_trace(n_jobs, lam)
    gp.render_forward(sc[0][2], _cam(sc[0][1], 128), 1 << 20, hero_c=4, max_depth=8)  # warm-up
    out = {"workload": f"c4 serving, {n_jobs} jobs over 8 resident scenes (Poisson {lam}/s, and saturated)",
           "scenes": [(n, int(sd.n_tris)) for n, sd, _ in sc],
           "photons": int(sum(j["photons"] for j in jobs)),
           "reference_style_tuning": REFERENCE_STYLE_TUNING}
    for mode, arrivals in (("poisson", [j["arrival"] for j in jobs]), ("saturated", [0.0] * len(jobs))):
        r = {"server": _run_server(lib, sc, jobs, arrivals),
             "server_admission_order": _run_server(lib, sc, jobs, arrivals, gp.abi.ORDER_ADMISSION),
             "server_one_live_job": _run_server(lib, sc, jobs, arrivals, gp.abi.ORDER_ADMISSION, 1),
             "plain_reference_style": _run_plain(lib, sc, jobs, arrivals, REFERENCE_STYLE_TUNING),
             "plain_tuned": _run_plain(lib, sc, jobs, arrivals)}
        r["throughput_ratio_vs_reference_style"] = (r["server"]["segments_per_s"] /
                                                    r["plain_reference_style"]["segments_per_s"])
        r["throughput_ratio_vs_tuned"] = r["server"]["segments_per_s"] / r["plain_tuned"]["segments_per_s"]
        out[mode] = r
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("config", choices=["c2", "c3", "c4"])
    ap.add_argument("--photons", type=int, default=1 << 28)
    ap.add_argument("--level", type=int, default=9)
    ap.add_argument("--jobs", type=int, default=60)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    a = ap.parse_args()
    if a.config == "c2":
        out = run_c2(a.photons, a.level)
    elif a.config == "c3":
        out = run_c3(autotune=not a.no_autotune)
    else:
        out = run_c4(a.jobs, full=not a.small)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
