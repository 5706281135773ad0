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
