"""The C++20 entry point (goodphotons_amd/host/gp_forward.hpp) called from a C++
program: single-shot renderForwardGpu equals a resident session traced in two
batches bit for bit, an unsupported camera model returns an empty film, a two-camera
SharedSession resumed from its downloaded totals equals the uninterrupted one, and
the film matches the oracle on the same scene and photons."""
import os
import subprocess

import numpy as np
import pytest

from goodphotons_amd import abi, scenes
from tests.parity import image_metrics

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "goodphotons_amd", "forward_demo")


def _read(path):
    raw = np.fromfile(path, dtype=np.float64)
    rx, ry = int(raw[0]), int(raw[1])
    e = raw[2:7]
    n = rx * ry
    xyz = raw[7:7 + 3 * n].reshape(n, 3)
    hits = raw[7 + 3 * n:7 + 4 * n]
    return xyz, hits, e


def _empty_cornell(res):
    b = scenes._Builder()
    scenes._cornell_shell(b, 0, 1, 2)
    b.quad((0.375, 0.999, 0.375), (0.625, 0.999, 0.375), (0.625, 0.999, 0.625), (0.375, 0.999, 0.625), 3)
    v, m = b.arrays()
    spd = abi.spectrum(abi.SPEC_BLACKBODY, 6500.0, 1.0, band=(380.0, 780.0))
    em = scenes.quad_emitter((0.375, 0.999, 0.375), (0.25, 0, 0), (0, 0, 0.25), (0, -1, 0), spd)
    return scenes.SceneData("empty_cornell", v, m, scenes._box_materials(), [em], scenes._std_camera(res))


def test_cpp_shim(tmp_path, orc):
    res, n = 64, 1 << 17
    out = str(tmp_path / "demo")
    r = subprocess.run([DEMO, out, str(res), str(n)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "unsupported_model_empty 1" in r.stdout
    # two-camera resident session: resumed from downloaded totals (seeded films) equals
    # uninterrupted bit for bit; camera 0's hit poll equals its downloaded hits; its
    # hits equal the single-camera render of the same photons
    assert "shared_session_resume_exact 1" in r.stdout
    assert "shared_session_hits0_equal 1" in r.stdout
    assert "shared_session_hits_equal_single 1" in r.stdout
    x1, h1, e1 = _read(out + "_single.bin")
    x2, h2, e2 = _read(out + "_session.bin")
    assert np.array_equal(x1, x2) and np.array_equal(h1, h2)
    assert np.allclose(e1, e2, rtol=1e-12)
    x5, h5, e5 = _read(out + "_shared0.bin")     # camera 0 of a two-camera shared flight
    assert np.array_equal(x1, x5) and np.array_equal(h1, h5)
    x4, h4, e4 = _read(out + "_resumed.bin")     # saved after batch 1, resumed elsewhere
    assert np.array_equal(x1, x4) and np.array_equal(h1, h4)
    assert np.allclose(e1, e4, rtol=1e-12)
    sd = _empty_cornell(res)
    cs = orc.scene(sd)
    cx, ch, ce, _ = orc.render(cs, sd.camera, n, hero_c=4, max_bounce=8, threads=8)
    m = image_metrics(x1, cx, n, n)
    assert m["rel_rmse_max"] < 0.01 and abs(m["mean_lum_ratio"] - 1) < 0.003
    assert abs(e1[0] / ce["emitted"] - 1) < 1e-6
