"""Test infrastructure only: the device BVH builder through the emulation build (child
process, started with the AddressSanitizer runtime preloaded by
tests/test_emulated_kernels.py). Builds small scenes on the "device" (a lone leaf, one
triangle more than a leaf holds, 300 triangles on one Morton code, a 5000-triangle
soup), probes closest hits against the oracle, and traces c0 with both builders.
Prints a JSON summary on the last line."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import goodphotons_amd as gp  # noqa: E402
import oracle  # noqa: E402
from goodphotons_amd import scenes  # noqa: E402
from tests import test_gpu_bvh_device as T  # noqa: E402
from tests.parity import seeded_rays  # noqa: E402

lib = gp.Library(os.path.join(ROOT, "tests", "emu", "_build", "libgp_emu.so"))
gp._LIB = lib
lib.set_tuning(queue_capacity=4096, grid_waves_per_cu=16)
orc = oracle.Oracle()
out = {}
for k in (1, 5):
    sd = T._only(T._tiny(0), k)
    T._check_probe(lib, orc, sd, seeded_rays(sd, 1000, 7), max_frac=0.2)
    out[sd.name] = "ok"
sd = T._tiny(300, dup=True)
T._check_probe(lib, orc, sd, seeded_rays(sd, 1000, 8), max_frac=1.0)
out[sd.name] = "ok"
sd = scenes.soup(5000, 7, res=32)
T._check_probe(lib, orc, sd, seeded_rays(sd, 1500, 4321))
out[sd.name] = "ok"
sd = scenes.cornell_c0(32)
films = []
for builder in ("host", "device"):
    gs = gp.GpuScene(sd, lib, builder=builder)
    x, h, e, st = gp.render_forward(gs, sd.camera, 1024, hero_c=4, max_depth=8, rr_depth=sd.rr_depth)
    films.append((x, h, st["segments"]))
(xa, ha, sa), (xb, hb, sb) = films
out["c0_films_identical"] = bool(sa == sb and np.array_equal(xa, xb) and np.array_equal(ha, hb))
print(json.dumps(out))
