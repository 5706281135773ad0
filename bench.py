"""Benchmark: traced segments/s of the forward light tracer on one MI355X.

Workload (the dispersive-caustics configuration): Cornell box + two diffuse boxes
+ 64x64-segment N-BK7 UV sphere + SF10 prism, 0.04 x 0.04 quad light (6500 K,
380-780 nm), pinhole 512x512, 2^26 photons per step, max depth 16, hero C = 4,
float32 transport, int64 fixed-point XYZ film. One step = one full render of the
2^26 photons (scene and film resident in HBM before the timed region).

Prints one JSON line. `value` = segments / wall time over the timed steps.
`roofline` is for the dominant kernel, from HIP events on the engine stream in a
profiled step of the same workload. `cpu_baseline` times the reference's own CPU
forward path (oracle/_ref, built from its sources; kind "reference"), or the oracle's
restatement of it where that build is absent (kind "port"), on a bounded sample of the
same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0

# Algorithmic bytes per work item of each kernel (structure-of-arrays records):
#   extend : read origin, direction, state (3 x 16 B), write hit (8 B)        per segment
#   shade  : read path (80 B) + hit (8 B) + triangle normal (16 B)            per segment
#            write survivor path (80 B)                                       per survivor
#            write connection (48 B)                                          per connection
#   emit   : write path (80 B)                                                per photon
#   connect: read connection (48 B)                                           per connection
#            film X, Y, Z, hits read-modify-write (4 x 8 B)                   per splat
EXTEND_B_PER_SEG = 56
SHADE_B_PER_SEG = 104
PATH_B = 80
CONN_B = 48
SPLAT_B = 32


def kernel_bytes(st):
    seg, ph, conn, spl = st["segments"], st["photons"], st["connections"], st["splats"]
    survivors = max(0, seg - ph)
    return {
        "k_extend": seg * EXTEND_B_PER_SEG,
        "k_shade": seg * SHADE_B_PER_SEG + survivors * PATH_B + conn * CONN_B,
        "k_emit": ph * PATH_B,
        "k_connect": conn * CONN_B + spl * SPLAT_B,
        "k_schedule": 0,
    }


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "c1_traffic.json")


def measured_traffic(kernel, tuning, photons):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes of this workload (scaled by the copy calibration), when they
    were taken at the same tuning and photon count; otherwise None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if photons != 1 << 26 or rec.get("tuning") != tuning or kernel not in rec.get("kernels", {}):
        return None
    return rec["kernels"][kernel]["hbm_bytes_per_dispatch"]


VALU_PROFILE = os.path.join(ROOT, "profiles", "c1_valu.json")
# wave64 VALU issue peak without packed FP32: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles
VALU_PEAK_G = 256 * 4 * 2.4 / 4


def measured_valu(kernel, tuning, photons, avg_ms):
    """VALU issue of `kernel` against the wave64 issue peak, from the committed
    rocprofv3 SQ_INSTS_VALU / SQ_THREAD_CYCLES_VALU pass of this workload (same
    tuning and photon count), divided by this run's live average launch time."""
    try:
        with open(VALU_PROFILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if photons != 1 << 26 or rec.get("tuning") != tuning or kernel not in rec.get("kernels", {}):
        return None
    k = rec["kernels"][kernel]
    rate = k["valu_insts_per_dispatch"] / (avg_ms * 1e-3) / 1e9
    return {"bound": "valu_issue", "achieved": rate, "peak": VALU_PEAK_G, "unit": "G wave-instructions/s",
            "frac": rate / VALU_PEAK_G, "lane_utilization": k["lane_utilization"],
            "source": "profiles/c1_valu.json"}


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(sd, target_s=10.0):
    """Time the CPU forward path on a bounded sample of the same workload.

    With the reference build present (oracle/_ref), the reference itself is timed
    (kind "reference"); the oracle traces the same photons to count segments exactly
    (its ray-query total must equal the reference's, a live bit-identity check).
    Otherwise the oracle is timed (kind "port")."""
    import oracle
    if not oracle.available("oracle"):
        return None
    orc = oracle.Oracle()
    cs = orc.scene(sd)
    T = cpu_threads()
    cam = sd.camera
    kw = dict(hero_c=sd.hero_c, max_bounce=sd.max_depth)
    ref = oracle.Reference() if oracle.available("ref") and sd.rr_depth == 0 else None
    timer = ref if ref is not None else orc
    ts = timer.scene(sd) if ref is not None else cs
    pilot = 1 << 16
    t0 = time.perf_counter()
    timer.render(ts, cam, pilot, seed_base=1 << 41, threads=T, **kw)
    dt = time.perf_counter() - t0
    n = int(min(1 << 25, max(1 << 16, pilot * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    _, _, _, tst = timer.render(ts, cam, n, seed_base=1 << 40, threads=T, **kw)
    dt = time.perf_counter() - t0
    _, _, _, st = orc.render(cs, cam, n, seed_base=1 << 40, threads=T, rr_depth=sd.rr_depth, **kw)
    out = {"value": st["segments"] / dt, "unit": "segments/s", "cores": T,
           "kind": "reference" if ref is not None else "port",
           "sample": f"{n} photons of the same workload (512x512, depth {sd.max_depth}, hero C={sd.hero_c}), "
                     f"{st['segments']} segments in {dt:.2f} s on {T} threads"}
    if ref is not None:
        out["reference_ray_queries"] = tst["ray_queries"]
        out["port_ray_queries"] = st["segments"] + st["connections"]
        out["bit_identical_ray_queries"] = tst["ray_queries"] == st["segments"] + st["connections"]
    return out


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime (the runtime's own
    copy kernel): the reference point that scales rocprof FETCH_SIZE/WRITE_SIZE."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--photons", type=int, default=1 << 26)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (rocprof scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (photon-range sharding "
              "plus one int64 film all-reduce is designed but not built); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import goodphotons_amd as gp
    from goodphotons_amd import scenes

    lib = gp.load()
    dev = lib.device_info()
    sd = scenes.caustics_c1(512)
    N = args.photons
    gs = gp.GpuScene(sd, lib)
    info = gs.info()
    if args.autotune:
        import ctypes as C
        t = gp.abi.Tuning()
        p = gp.trace_params(N, hero_c=sd.hero_c, max_depth=sd.max_depth, rr_depth=sd.rr_depth)
        lib.check(lib.lib.gp_autotune(gs.h, C.byref(sd.camera), C.byref(p), C.byref(t)))
    film = gp.Film(512, 512, max_photons=N, lib=lib)

    def step(seed_base):
        film.reset()
        e, st = gp.trace(gs, sd.camera, film, N, seed_base=seed_base, hero_c=sd.hero_c,
                         max_depth=sd.max_depth, rr_depth=sd.rr_depth)
        return st.as_dict()

    for w in range(args.warmup):
        step(w * N)
    segs = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = step((args.warmup + k) * N)
        segs += st["segments"]
    dt = time.perf_counter() - t0
    value = segs / dt

    # Profiled step: per-kernel HIP-event times on the engine stream.
    tun = lib.tuning()
    lib.set_tuning(profile=1)
    try:
        pst = step(0)
    finally:
        lib.set_tuning(profile=0)
    kb = kernel_bytes(pst)
    kms = dict(zip(gp.abi.KERNEL_NAMES, pst["kernel_ms"]))
    kl = dict(zip(gp.abi.KERNEL_NAMES, pst["kernel_launches"]))
    dom = max(kms, key=kms.get)
    avg_ms = kms[dom] / max(1, kl[dom])
    bytes_per_launch = kb[dom] / max(1, kl[dom])
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    total_bytes = sum(kb.values())
    total_ms = sum(kms.values())
    traffic = measured_traffic(dom, {k: getattr(tun, k) for k, _ in tun._fields_}, N)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                "traffic_source": "profiles/c1_traffic.json" if traffic is not None else None,
                "avg_launch_ms": avg_ms, "launches": kl[dom],
                "algorithmic_bytes_per_launch": bytes_per_launch,
                "valu": measured_valu(dom, {k: getattr(tun, k) for k, _ in tun._fields_}, N, avg_ms),
                "kernel_ms": kms, "pipeline_bytes_per_segment": total_bytes / max(1, pst["segments"]),
                "pipeline_GBps": total_bytes / (total_ms * 1e-3) / 1e9}

    cpu = None if args.no_cpu else cpu_baseline(sd)
    out = {
        "metric": "traced segments/s at matched 1% image RMSE",
        "value": value, "unit": "segments/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c1 dispersive caustics: Cornell box + 2 boxes + 64x64-segment N-BK7 "
                               "UV sphere + SF10 prism, 0.04x0.04 quad light 6500K, 512x512 pinhole, "
                               "2^26 photons/step, depth 16, hero C=4, int64 XYZ film",
                   "photons_per_step": N, "tris": info["n_tris"], "bvh_depth": info["bvh_depth"],
                   "segments_per_step": segs / args.steps,
                   "device": f"{dev['name'] or 'AMD GPU'} ({dev['arch']}, {dev['compute_units']} CUs)",
                   "tuning": {k: getattr(tun, k) for k, _ in tun._fields_},
                   "matched_rmse": "GPU and CPU trace the same photons (same per-photon PCG "
                                   "streams), so equal photon count is matched RMSE"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
