"""A caller's own binary BVH (gp_scene_create_ex with GP_BVH_PREBUILT, the reference's
Scene::bvh handed over as built) is checked as it is read, before any device work, so
these run on the CPU: a well-formed tree passes every check (the call then stops at the
missing device), and each malformed tree is refused with GP_ERR_INVALID and the
reason."""
import ctypes as C

import numpy as np
import pytest

from goodphotons_amd import abi, scenes

GP_ERR_INVALID, GP_ERR_NO_DEVICE = 1, 3

# two triangles side by side (x in [0, 1] and [2, 3]) and a light quad above them
TRIS = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0],
                 [2, 0, 0, 3, 0, 0, 2, 1, 0]], np.float32)


def _node(lo, hi, left=-1, right=-1, first=0, count=0):
    n = abi.BvhNode()
    for a in range(3):
        n.lo[a], n.hi[a] = lo[a], hi[a]
    n.left, n.right, n.first, n.count = left, right, first, count
    return n


def _good():
    e = 1e-6
    return [_node((-e, -e, -e), (3 + e, 1 + e, e), left=1, right=2),
            _node((-e, -e, -e), (1 + e, 1 + e, e), first=0, count=1),
            _node((2 - e, -e, -e), (3 + e, 1 + e, e), first=1, count=1)], [0, 1]


def _create(nodes, prim):
    lib = C.CDLL(abi.LIB_PATH)
    lib.gp_scene_create_ex.argtypes = [C.POINTER(abi.SceneDesc), C.c_int32, C.POINTER(C.c_void_p)]
    lib.gp_last_error.restype = C.c_char_p
    sd = scenes.SceneData("two", TRIS.copy(), np.zeros(2, np.int32),
                          [scenes.diffuse(abi.spectrum(abi.SPEC_CONST, 0.5))],
                          [scenes.quad_emitter((0, 2, 0), (1, 0, 0), (0, 0, 1), (0, -1, 0),
                                               abi.spectrum(abi.SPEC_CONST, 1.0))],
                          scenes._std_camera(8))
    d = sd.desc()
    arr = (abi.BvhNode * len(nodes))(*nodes)
    idx = np.ascontiguousarray(prim, np.int32)
    d.bvh_nodes = C.cast(arr, C.POINTER(abi.BvhNode))
    d.n_bvh_nodes = len(nodes)
    d.bvh_prim_index = idx.ctypes.data_as(C.POINTER(C.c_int32))
    d.n_bvh_prim_index = len(idx)
    h = C.c_void_p()
    st = lib.gp_scene_create_ex(C.byref(d), abi.BVH_PREBUILT, C.byref(h))
    return st, lib.gp_last_error().decode()


def test_well_formed_tree_passes_the_checks():
    nodes, prim = _good()
    st, msg = _create(nodes, prim)
    assert st == GP_ERR_NO_DEVICE, msg


def _broken(kind):
    nodes, prim = _good()
    if kind == "child outside parent":
        nodes[2].hi[0] = 4.0
    elif kind == "triangle outside leaf":
        nodes[1].hi[0] = 0.5
        nodes[0].hi[0] = 3 + 1e-6
    elif kind == "triangle in no leaf":
        prim = [0, 0]           # both leaves hold triangle 0 (the second leaf's box grown to it)
        nodes[2].lo[0] = -1e-6
    elif kind == "node reached twice":
        nodes[0].right = 1
    elif kind == "child out of range":
        nodes[0].right = 7
    elif kind == "leaf too large":
        nodes[1].count = 5
    elif kind == "leaf range outside index":
        nodes[2].first = 1
        nodes[2].count = 2
    elif kind == "prim out of range":
        prim = [0, 9]
    return nodes, prim


@pytest.mark.parametrize("kind,reason", [
    ("child outside parent", "a child box is not inside its parent's"),
    ("triangle outside leaf", "a leaf box does not contain its triangle"),
    ("triangle in no leaf", "a triangle is in no leaf"),
    ("node reached twice", "a node is reached twice"),
    ("child out of range", "child index out of range"),
    ("leaf too large", "a leaf holds more than 4 primitives"),
    ("leaf range outside index", "leaf range outside the primitive index"),
    ("prim out of range", "primitive index out of range"),
])
def test_malformed_trees_are_refused(kind, reason):
    st, msg = _create(*_broken(kind))
    assert st == GP_ERR_INVALID and reason in msg, msg


def test_prebuilt_without_a_tree_is_refused():
    st, msg = _create([], [])
    assert st == GP_ERR_INVALID and "bvh_nodes" in msg
