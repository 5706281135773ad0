"""Test infrastructure only: one trace through the emulation build (child process,
started with the AddressSanitizer runtime preloaded by tests/test_emulated_kernels.py).
Prints the stats and the film as JSON on the last line."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import goodphotons_amd as gp  # noqa: E402
from goodphotons_amd import scenes  # noqa: E402

name, cap, n, res = sys.argv[1], int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4])
lds = int(sys.argv[5]) if len(sys.argv) > 5 else 16
ray_sort = int(sys.argv[6]) if len(sys.argv) > 6 else 0
lib = gp.Library(os.path.join(ROOT, "tests", "emu", "_build", "libgp_emu.so"))
lib.set_tuning(queue_capacity=cap, lds_stack=lds, grid_waves_per_cu=16, ray_sort=ray_sort)
sd = {"c0": scenes.cornell_c0, "c1": scenes.caustics_c1, "glossy": scenes.glossy_box,
      "mirror": scenes.mirror_box}[name](res)
gs = gp.GpuScene(sd, lib)
x, h, e, st = gp.render_forward(gs, sd.camera, n, hero_c=4, max_depth=8)
print(json.dumps({"stats": st, "energy": e, "xyz": x.tolist(), "hits": h.tolist()}))
