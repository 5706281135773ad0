"""The device-built linear BVH (gp_scene_create_ex, GP_BVH_DEVICE_LBVH): closest-hit
primitive ids against the oracle to the same bar as the host-built tree (bit-exact
except rays the oracle marks ambiguous), films bit-identical to the host-built tree's
on the same photons, edge cases of the Karras construction (a lone leaf, one more
triangle than a leaf holds, thousands of triangles sharing one Morton code), and the
5.24 M-triangle c2 mesh built in well under a second."""
import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import abi, scenes
from tests.parity import check_prim_parity, seeded_rays

pytestmark = pytest.mark.gpu


def _check_probe(gplib, orc, sd, rays, max_frac=0.1):
    gs = gp.GpuScene(sd, gplib, builder="device")
    info = gs.info()
    cs = orc.scene(sd)
    gprim, gt = gs.closest_hit(rays, tmin=1e-6)
    cprim, ct = cs.closest_hit(rays.astype(np.float64), tmin=1e-6)
    print(f"[lbvh] {sd.name} tris={sd.n_tris} depth={info['bvh_depth']} build={info['build_ms']:.1f} ms",
          flush=True)
    check_prim_parity(orc, cs, rays, gprim, gt, cprim, ct, "lbvh_" + sd.name, max_tie_frac=max_frac)
    return gs, info


@pytest.mark.parametrize("name,make,n", [
    ("c0", lambda: scenes.cornell_c0(64), 100000),
    ("c1", lambda: scenes.caustics_c1(64), 100000),
    ("c2_l5", lambda: scenes.icosphere_c2(level=5, res=64), 50000),
    ("soup20k", lambda: scenes.soup(20000, 7, res=64), 50000),
], ids=["c0", "c1", "c2_l5", "soup20k"])
def test_device_bvh_prim_ids(gplib, orc, name, make, n):
    sd = make()
    _check_probe(gplib, orc, sd, seeded_rays(sd, n, 4321))


def _tiny(n_tris, dup=False):
    b = scenes._Builder()
    scenes._cornell_shell(b, 0, 1, 2)       # 10 wall triangles
    rng = np.random.default_rng(3)
    v0 = np.array([0.4, 0.3, 0.5])
    for k in range(n_tris):
        o = v0 if dup else v0 + rng.uniform(-0.2, 0.2, 3)
        b.tri(tuple(o), tuple(o + [0.1, 0.0, 0.02]), tuple(o + [0.0, 0.1, 0.01]), 0)
    v, m = b.arrays()
    spd = abi.spectrum(abi.SPEC_BLACKBODY, 6500.0, 1.0, band=(380.0, 780.0))
    em = scenes.quad_emitter((0.375, 0.999, 0.375), (0.25, 0, 0), (0, 0, 0.25), (0, -1, 0), spd)
    return scenes.SceneData(f"tiny{n_tris}{'dup' if dup else ''}", v, m, scenes._box_materials(), [em],
                            scenes._std_camera(32))


def _only(sd, k):
    """The same scene cut down to its first k triangles."""
    return scenes.SceneData(sd.name + f"_first{k}", sd.verts[:k].copy(), sd.mat_ids[:k].copy(),
                            sd.materials, sd.emitters, sd.camera)


@pytest.mark.parametrize("k", [1, 4, 5])
def test_device_bvh_small_scenes(gplib, orc, k):
    sd = _only(_tiny(0), k)
    _check_probe(gplib, orc, sd, seeded_rays(sd, 20000, 7), max_frac=0.2)


def test_device_bvh_duplicate_morton_codes(gplib, orc):
    sd = _tiny(3000, dup=True)          # 3000 identical triangles: one Morton code
    gs, info = _check_probe(gplib, orc, sd, seeded_rays(sd, 20000, 8), max_frac=1.0)
    assert info["bvh_depth"] < 40


@pytest.mark.parametrize("name,make,n,depth", [
    ("c0", lambda: scenes.cornell_c0(64), 1 << 17, 8),
    ("c1", lambda: scenes.caustics_c1(64), 1 << 17, 16),
])
def test_device_bvh_film_equals_host_bvh(gplib, name, make, n, depth):
    sd = make()
    films = []
    for builder in ("host", "device"):
        gs = gp.GpuScene(sd, gplib, builder=builder)
        x, h, e, st = gp.render_forward(gs, sd.camera, n, hero_c=4, max_depth=depth, rr_depth=sd.rr_depth)
        films.append((x, h, st["segments"]))
    (xa, ha, sa), (xb, hb, sb) = films
    print(f"[lbvh] {name} film: segments host {sa} device {sb}, identical {np.array_equal(xa, xb)}", flush=True)
    assert sa == sb and np.array_equal(xa, xb) and np.array_equal(ha, hb)


def test_device_bvh_c2_full_scale(gplib, orc):
    sd = scenes.icosphere_c2(level=9, res=64)
    rng = np.random.default_rng(11)
    n = 100000
    o = np.array([0.5, 0.31, 0.5]) + rng.normal(size=(n, 3)) * 0.6
    tgt = np.array([0.5, 0.31, 0.5]) + rng.normal(size=(n, 3)) * 0.15
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    gs, info = _check_probe(gplib, orc, sd, rays, max_frac=0.001)
    assert info["build_ms"] < 1000
