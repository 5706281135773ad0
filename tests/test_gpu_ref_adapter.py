"""renderForwardCuda's replacement through the drop-in adapter on the device: scenes
built with the reference's own API (oracle/_ref/libref_adapter.so: the adapter
compiled against the reference's headers), rendered by gpref::renderForward on the
HIP library with the scene's own BVH handed over (GP_BVH_PREBUILT) and with the
library's SAH tree, against the reference's CPU forward path on the same scene and
photons: per-channel relative RMSE < 1%, mean luminance within 0.3%, equal emitted
energy; the two trees give the same film up to float rounding."""
import numpy as np
import pytest

import oracle
from tests.parity import host_threads, image_metrics

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle.available("adapter"),
                                 reason="oracle/_ref/libref_adapter.so not built (reference sources absent)")]

GP_BVH_HOST_SAH, GP_BVH_PREBUILT = 0, 2


@pytest.mark.parametrize("which,hero_c", [(0, 4), (1, 4), (1, 1)])
def test_adapter_render_matches_reference(gplib, which, hero_c):
    adp = oracle.RefAdapter()
    s = adp.scene(which)
    res, n = 64, 1 << 18
    print(f"[adapter] scene {which} hero {hero_c} n={n}", flush=True)
    gx, gh, ge = adp.render_gpu(s, res, n, hero_c=hero_c, depth=8, bvh=GP_BVH_PREBUILT)
    sx, sh, se = adp.render_gpu(s, res, n, hero_c=hero_c, depth=8, bvh=GP_BVH_HOST_SAH)
    rx, rh, re = adp.render_reference(s, res, n, hero_c=hero_c, depth=8, threads=host_threads())
    m = image_metrics(gx, rx, n, n, res=(res, res))
    t = image_metrics(gx, sx, n, n, res=(res, res))
    print(f"[adapter] vs reference {m}; prebuilt vs own tree {t}", flush=True)
    assert gx.any() and m["rel_rmse_max"] < 0.01 and abs(m["mean_lum_ratio"] - 1) < 0.003
    assert ge["emitted"] == pytest.approx(re["emitted"], rel=1e-6)
    assert abs(gh.sum() / rh.sum() - 1) < 2e-3
    assert t["rel_rmse_max"] < 1e-4 and ge["emitted"] == se["emitted"]


@pytest.mark.parametrize("which", [0, 1])
def test_adapter_shared_session_resumes_exactly(gplib, which):
    # sharedForwardGpuBegin / Batch / Hits0 / Download / End with the reference's types:
    # two cameras, 4 batches; stopped after 2, downloaded, and resumed as seedFilms /
    # seedEnergy, the totals equal the uninterrupted session's bit for bit (the seed
    # sums round back to the same fixed-point words); the camera-0 poll equals the
    # downloaded hit counts; camera 0 matches the single-camera render of the same
    # photons.
    adp = oracle.RefAdapter()
    s = adp.scene(which)
    res, res2, n = 48, 40, 1 << 16
    print(f"[adapter session] scene {which}", flush=True)
    films, h0, e, rfilms, re = adp.shared_session(s, res, res2, n, 4, 2, hero_c=4, depth=8)
    for (x, h), (rx, rh) in zip(films, rfilms):
        assert h.sum() > 0 and np.array_equal(x, rx) and np.array_equal(h, rh)
    assert np.array_equal(h0, films[0][1])
    assert np.allclose(e, re, rtol=1e-12) and e[0] > 0
    gx, gh, ge = adp.render_gpu(s, res, 4 * n, hero_c=4, depth=8)
    assert np.array_equal(gh, films[0][1])
    assert image_metrics(films[0][0], gx, 4 * n, 4 * n)["rel_rmse_max"] < 1e-5
