"""The library's flat-mirror plane builder (gp_mirror_planes, host only) against the
reference's planes (golden, written by the reference's Scene::buildMirrorPlanes) and
the oracle's restatement, bit for bit. No GPU needed."""
import math

import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import abi, scenes
from tests.conftest import golden


@pytest.fixture(scope="module")
def lib():
    return gp.Library()


def test_planes_equal_reference_golden(lib):
    g = golden("mirror_planes.npz")["planes"]
    p = lib.mirror_planes(scenes.mirror_box(32))
    assert p.shape == (7, 10) and np.array_equal(p, g)


def test_no_mirrors_no_planes(lib):
    assert lib.mirror_planes(scenes.cornell_c0(16)).shape == (0, 10)
    assert lib.mirror_planes(scenes.caustics_c1(16)).shape == (0, 10)


def _mirror_field(n_quads, seed, split=True):
    """Mirror quads in random orientations and sizes (all areas distinct), some split
    into coplanar halves that must share a plane, plus a degenerate triangle."""
    rng = np.random.default_rng(seed)
    s = scenes.cornell_c0(16)
    mats = list(s.materials) + [scenes.mirror(abi.spectrum(abi.SPEC_CONST, 0.9))]
    mi = len(mats) - 1
    tris = [s.verts]
    ids = [s.mat_ids]
    for k in range(n_quads):
        c = rng.uniform(0.1, 0.9, 3)
        a, b = rng.uniform(0, 2 * math.pi), rng.uniform(0, math.pi)
        nrm = np.array([math.sin(b) * math.cos(a), math.sin(b) * math.sin(a), math.cos(b)])
        t1 = np.cross(nrm, [0.3, 0.5, 0.81])
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(nrm, t1)
        h = 0.01 + 0.002 * k   # distinct areas
        q = [c - h * t1 - h * t2, c + h * t1 - h * t2, c + h * t1 + h * t2, c - h * t1 + h * t2]
        quad = [q[0], q[1], q[2], q[0], q[2], q[3]] if split else [q[0], q[1], q[2]]
        tris.append(np.array(quad, np.float32).reshape(-1, 9))
        ids.append(np.full(len(quad) // 3, mi, np.int32))
    tris.append(np.array([[0.2, 0.2, 0.2, 0.2, 0.2, 0.2, 0.4, 0.5, 0.6]], np.float32))  # zero area
    ids.append(np.array([mi], np.int32))
    return scenes.SceneData("mirror_field", np.concatenate(tris), np.concatenate(ids), mats,
                            list(s.emitters), s.camera)


@pytest.mark.parametrize("n_quads", [5, 40, 80])
def test_planes_equal_oracle(lib, orc, n_quads):
    # 80 planes exceed the 64-plane cap: the largest areas are kept
    sd = _mirror_field(n_quads, seed=n_quads)
    p = lib.mirror_planes(sd)
    o = orc.scene(sd).mirror_planes()
    # the oracle keeps the degenerate triangle's NaN plane (as the reference does);
    # the library drops it: it can never pass the connection's same-side test
    o = o[np.all(np.isfinite(o), axis=1)]
    assert np.array_equal(p, o)
    # coplanar halves mostly share a plane (a split can fall either side of a key's
    # rounding boundary: then both sides keep two entries)
    assert min(n_quads, 64) <= p.shape[0] <= min(64, 2 * n_quads)
    if n_quads > 64:
        assert p.shape[0] == 64


def test_bad_arguments(lib):
    import ctypes as C
    assert lib.lib.gp_mirror_planes(None, None, 0) == -1   # -GP_ERR_INVALID
    d = scenes.mirror_box(16).desc()
    assert lib.lib.gp_mirror_planes(C.byref(d), None, 3) == -1   # -GP_ERR_INVALID
    assert lib.lib.gp_mirror_planes(C.byref(d), None, 0) == 7
