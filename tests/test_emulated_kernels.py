"""The unchanged kernel sources run on the CPU under AddressSanitizer (tests/emu: each
GPU lane an emulated thread, device memory filled with garbage at allocation), so every
memory access of the device code is bounds-checked without a GPU.

- A small trace with lanes preempted inside their loops (so lanes of a block interleave
  as waves do) traces exactly the oracle's photons: equal segment, connection and splat
  counts and the same film up to float rounding.
- The flat-mirror connection (k_mirror) with interleaved lanes: the configuration in
  which its old shared LDS stack made lanes pop garbage nodes; it must trace the
  oracle's photons, with the traversal stack in LDS and with an 8-entry LDS stack that
  spills to global memory.
- Regression for the splat-table race in k_connect: at 128^2 with a 2^14-path queue on
  a 4-workgroup grid, the old table re-insertion (no barrier after it) indexed the LDS
  table with -1 and sent a splat to a garbage film address here; it must run clean.
- Ray binning (the counting sort of each iteration's paths into a third queue, and
  the closest-hit and shading kernels reading it): bit-identical to the unbinned trace.
- The device BVH builder (linear BVH and its collapse into four-wide nodes) with the
  four-wide traversal: closest-hit prim ids against the oracle on edge-case scenes and
  a soup, and c0 films identical with the host-built and the device-built tree.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from goodphotons_amd import scenes
from tests.emu import build as emu
from tests.parity import image_metrics

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(not emu.available(), reason="clang++ with AddressSanitizer not found")


def _env(cus, preempt=0):
    env = dict(os.environ)
    env.update({"LD_PRELOAD": emu.asan_runtime(), "ASAN_OPTIONS": "detect_leaks=0",
                "EMU_CUS": str(cus), "EMU_PREEMPT": str(preempt)})
    return env


def _trace(name, cap, n, res, lds=16, ray_sort=0):
    return [sys.executable, os.path.join(ROOT, "tests", "emu", "run_trace.py"), name, str(cap), str(n), str(res),
            str(lds), str(ray_sort)]


# Every emulated run is one mostly single-threaded child process (one lane runs at a
# time), so the module starts them all at once and each test waits for its own.
CASES = {
    "c1_interleaved": (_trace("c1", 4096, 4096, 32), _env(2, preempt=3)),
    "c1_binned": (_trace("c1", 4096, 4096, 32, 12, ray_sort=1), _env(2, preempt=3)),
    "mirror_lds": (_trace("mirror", 4096, 2048, 32, 16), _env(2, preempt=3)),
    "mirror_spill": (_trace("mirror", 4096, 2048, 32, 8), _env(2, preempt=3)),
    "splat_flush": (_trace("c0", 1 << 14, 1 << 14, 128), _env(1)),
    "device_bvh": ([sys.executable, os.path.join(ROOT, "tests", "emu", "run_lbvh.py")], _env(2)),
}


@pytest.fixture(scope="module")
def runs():
    emu.build()
    procs = {k: subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env)
             for k, (cmd, env) in CASES.items()}
    yield procs
    for p in procs.values():
        if p.poll() is None:
            p.kill()
            p.wait()


def _result(runs, key):
    out, err = runs[key].communicate(timeout=900)
    assert runs[key].returncode == 0, err[-3000:]
    return json.loads(out.strip().splitlines()[-1])


def _film(runs, key):
    out = _result(runs, key)
    return np.array(out["xyz"]), np.array(out["hits"]), out["energy"], out["stats"]


def _oracle(orc, name, n, res):
    sd = {"c0": scenes.cornell_c0, "c1": scenes.caustics_c1, "glossy": scenes.glossy_box,
          "mirror": scenes.mirror_box}[name](res)
    return orc.render(orc.scene(sd), sd.camera, n, hero_c=4, max_bounce=8, threads=8)


def test_interleaved_lanes_trace_the_oracles_photons(runs, orc):
    x, h, e, st = _film(runs, "c1_interleaved")
    cx, ch, ce, cst = _oracle(orc, "c1", 4096, 32)
    assert (st["segments"], st["connections"], st["splats"]) == (cst["segments"], cst["connections"], cst["splats"])
    m = image_metrics(x, cx, 4096, 4096, res=(32, 32))
    assert m["rel_rmse_max"] < 1e-5 and np.array_equal(h, ch)


def test_ray_binning_changes_nothing(runs, orc):
    # the binning kernels (count, scan, scatter into the third queue) and the binned
    # reads of extend/shade, lanes interleaved, with a 12-entry closest-hit LDS stack
    # (the any-hit kernel then keeps 8): the same photons, bit for bit
    x, h, e, st = _film(runs, "c1_binned")
    px, ph, pe, pst = _film(runs, "c1_interleaved")
    assert np.array_equal(x, px) and np.array_equal(h, ph) and e == pe
    assert (st["segments"], st["connections"], st["splats"]) == (pst["segments"], pst["connections"], pst["splats"])
    cx, ch, ce, cst = _oracle(orc, "c1", 4096, 32)
    assert st["segments"] == cst["segments"] and np.array_equal(h, ch)


@pytest.mark.parametrize("key", ["mirror_lds", "mirror_spill"], ids=["lds_stack", "spilling_stack"])
def test_mirror_connection_with_interleaved_lanes(runs, orc, key):
    x, h, e, st = _film(runs, key)
    cx, ch, ce, cst = _oracle(orc, "mirror", 2048, 32)
    assert (st["segments"], st["connections"], st["splats"]) == (cst["segments"], cst["connections"], cst["splats"])
    m = image_metrics(x, cx, 2048, 2048, res=(32, 32))
    assert m["rel_rmse_max"] < 1e-5 and np.array_equal(h, ch)


def test_splat_table_flush_under_load(runs, orc):
    x, h, e, st = _film(runs, "splat_flush")
    cx, ch, ce, cst = _oracle(orc, "c0", 1 << 14, 128)
    assert abs(st["segments"] / cst["segments"] - 1) < 1e-3
    assert abs(st["splats"] / cst["splats"] - 1) < 1e-3
    m = image_metrics(x, cx, 1 << 14, 1 << 14, res=(128, 128))
    assert m["rel_rmse_max"] < 1e-3


def test_device_bvh_builder_and_wide_traversal(runs):
    out = _result(runs, "device_bvh")
    assert out["c0_films_identical"] and len(out) == 5
