"""Shared helpers for the GPU-vs-oracle parity tests, smoke() and bench.py."""
from __future__ import annotations

import numpy as np

YBAR_INTEGRAL = 106.94618833597131   # sum of the fitted ybar over 360..830 nm at 1 nm


def seeded_rays(sd, n, seed):
    """Rays from points around the scene's bounding box in uniform directions
    (float32-exact), the -checkbvh pattern (src/main.cpp:448-473)."""
    rng = np.random.default_rng(seed)
    v = sd.verts.reshape(-1, 3).astype(np.float64)
    lo, hi = v.min(0), v.max(0)
    o = lo + (hi - lo) * rng.uniform(-0.1, 1.1, size=(n, 3))
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    return np.concatenate([o, d], 1).astype(np.float32)


def image_metrics(gpu_xyz, ref_xyz, n_gpu, n_ref, box=8, res=None):
    """Relative-RMSE metrics between two films (un-normalised XYZ sums).

    Films are normalised by 1/(N * integral(ybar)) as the reference's filmToLinear
    does (src/main.cpp:10282-10294). Per channel: sqrt(sum (G-R)^2) / sqrt(sum R^2)
    over pixels lit in R; mean luminance ratio of Y; and the relative RMS of the
    luminance after a box x box downsample (tools/pfmcmp.py:40-63)."""
    G = gpu_xyz / (n_gpu * YBAR_INTEGRAL)
    R = ref_xyz / (n_ref * YBAR_INTEGRAL)
    lit = R[:, 1] > 0
    rel = []
    for c in range(3):
        num = np.sqrt(np.sum((G[lit, c] - R[lit, c]) ** 2))
        den = np.sqrt(np.sum(R[lit, c] ** 2))
        rel.append(float(num / den) if den > 0 else 0.0)
    lum_ratio = float(G[:, 1].mean() / R[:, 1].mean())
    out = {"rel_rmse_xyz": rel, "rel_rmse_max": max(rel), "mean_lum_ratio": lum_ratio}
    if res is not None and res[0] % box == 0 and res[1] % box == 0:
        gy = G[:, 1].reshape(res[1] // box, box, res[0] // box, box).mean(axis=(1, 3))
        ry = R[:, 1].reshape(res[1] // box, box, res[0] // box, box).mean(axis=(1, 3))
        m = ry > 1e-30
        out["rel_rms_box"] = float(np.sqrt(np.mean((gy[m] / ry[m] - 1.0) ** 2)))
    return out


# Closest-hit parity rule. Rays whose two nearest hits lie within TIE_REL relative
# distance are excluded from prim-id parity (either triangle is a correct answer).
# A float32 test can also decide differently from the double one where the ray
# grazes a triangle's boundary; such rays are counted apart and bounded by
# GRAZE_BOUND (a fraction of the probed rays), never excused silently.
TIE_REL = 1e-6
GRAZE_MARGIN = 1e-5       # relative margin on the scaled edge functions
GRAZE_BOUND = 1e-4


def classify_prim_mismatches(orc, cs, rays, gprim, gt, cprim, ct, tmin=1e-6):
    """Classify every ray whose GPU prim id differs from the oracle's, from the
    oracle's brute-force ambiguity report (or_closest_hit_ambiguity) with the GPU's
    triangle as the queried one:

    near_tie    both sides hit, and the GPU's triangle is hit (edges grown by the
                margin) at a distance within TIE_REL of the oracle's winner: the two
                nearest hits lie within 1e-6 relative t;
    edge        both sides hit, and a grazed triangle (hit with its edges grown by the
                margin, missed with them shrunk) explains the difference: the oracle's
                winner is grazed and the GPU returned the nearest hit behind it (no
                robust hit in front of the GPU's), or the GPU hit a grazed triangle in
                front of the oracle's winner;
    miss_graze  the same with one side missing: the oracle misses and the GPU hit a
                grazed triangle, or the GPU misses, the oracle's winner is grazed and
                nothing is robustly hit;
    unexplained everything else (a real mismatch).

    The GPU's distance must agree with its triangle's double distance (float32
    tolerance) for any of the first three. Returns (counts, labels)."""
    diff = np.nonzero(gprim != cprim)[0]
    labels = np.array(["unexplained"] * len(diff), dtype=object)
    if len(diff):
        a = orc.closest_hit_ambiguity(cs, np.asarray(rays)[diff].astype(np.float64), gprim[diff],
                                      tmin=tmin, rel_margin=GRAZE_MARGIN)
        # (on an exact tie the brute-force winner may be another of the tied triangles
        # than the BVH's, at the same distance)
        assert np.array_equal(a["t"], ct[diff])
        g, c = gprim[diff], cprim[diff]
        t, tq, trob = a["t"], a["tq"], a["t_robust"]
        gtd = np.asarray(gt, np.float64)[diff]
        with np.errstate(invalid="ignore"):   # inf - inf on misses: compares False
            t_ok = np.isfinite(tq) & (np.abs(gtd - tq) <= 2e-6 * (1.0 + np.where(np.isfinite(tq), tq, 0.0)))
            near = (c >= 0) & (g >= 0) & t_ok & (np.abs(tq - t) <= TIE_REL * np.abs(t))
        behind = a["edge"] & (g >= 0) & t_ok & (tq >= t * (1 - TIE_REL)) & (tq <= trob * (1 + TIE_REL) + 2e-6)
        front = (g >= 0) & a["q_graze"] & t_ok & (tq <= t * (1 + TIE_REL))
        gpu_miss = a["edge"] & (g < 0) & ~np.isfinite(trob)
        edge = ~near & (c >= 0) & (g >= 0) & (behind | front)
        miss = ~near & ~edge & (((c < 0) & front) | gpu_miss)
        labels[near] = "near_tie"
        labels[edge] = "edge"
        labels[miss] = "miss_graze"
    counts = {k: int((labels == k).sum()) for k in ("near_tie", "edge", "miss_graze", "unexplained")}
    counts["rays"] = int(len(gprim))
    return counts, labels


def check_prim_parity(orc, cs, rays, gprim, gt, cprim, ct, name, max_tie_frac=0.1, tmin=1e-6):
    """The closest-hit parity bar: no unexplained prim-id mismatch, grazes within
    GRAZE_BOUND of the rays, near-ties within max_tie_frac (a scene property: e.g.
    boxes standing on the floor are exact coplanar ties), and equal distances wherever
    the prim ids agree."""
    counts, _ = classify_prim_mismatches(orc, cs, rays, gprim, gt, cprim, ct, tmin=tmin)
    n = counts["rays"]
    print(f"[probe] {name}: rays={n} near_tie={counts['near_tie']} edge={counts['edge']} "
          f"miss_graze={counts['miss_graze']} unexplained={counts['unexplained']}", flush=True)
    assert counts["unexplained"] == 0, f"{counts['unexplained']} unexplained prim-id mismatches"
    assert counts["edge"] + counts["miss_graze"] <= GRAZE_BOUND * n, counts
    assert counts["near_tie"] <= max_tie_frac * n, counts
    both = (gprim == cprim) & (cprim >= 0)
    assert np.all(np.abs(gt[both] - ct[both]) <= 2e-6 * (1.0 + ct[both]))
    return counts


def host_threads():
    """CPU threads for oracle runs: the cores this process may use, capped by the
    host's per-GPU CPU share when OMP_NUM_THREADS sets one."""
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


SAMPLED_LO, SAMPLED_HI, SAMPLED_N = 360.5, 829.5, 470


def to_sampled(sd, orc):
    """The same scene with every material and emitter spectrum replaced by its values at
    the 470 emission-bin centres 360.5 .. 829.5 nm (SPEC_SAMPLED, values from the
    oracle's analytic evaluation), the form in which the reference's Spectrum
    functions reach the core through the adapter."""
    from goodphotons_amd import abi
    lam = SAMPLED_LO + np.arange(SAMPLED_N, dtype=np.float64)
    pool = []

    def conv(spec):
        s = abi.Spectrum()
        s.kind = abi.SPEC_SAMPLED
        s.table = SAMPLED_N * len(pool)
        s.p[0], s.p[1], s.p[2] = SAMPLED_LO, SAMPLED_HI, SAMPLED_N
        pool.append(orc.spectrum_eval(spec, lam))
        return s

    mats = []
    for m in sd.materials:
        m2 = type(m).from_buffer_copy(m)
        m2.reflect = conv(m.reflect)
        m2.ior = conv(m.ior)
        mats.append(m2)
    ems = []
    for e in sd.emitters:
        e2 = type(e).from_buffer_copy(e)
        e2.spd = conv(e.spd)
        ems.append(e2)
    out = sd.with_camera_res(sd.camera.res_x, sd.camera.res_y)
    out.name = sd.name + "_sampled"
    out.materials, out.emitters = mats, ems
    out.spectrum_samples = np.concatenate(pool)
    return out
