"""The oracle's closest-hit ambiguity report (or_closest_hit_ambiguity) and the
prim-id mismatch classification built on it, pinned on constructed rays whose answers
the geometry fixes (CPU only).

Rays travel along -z from z = 1 onto triangles in the z = 0 plane, so every hit is at
t = 1 exactly. For the triangle (0,0,0), (1,0,0), (0,1,0) and a ray through (0.25, y0)
the sheared-frame edge functions are U = 0.75 - y0, V = 0.25, W = y0 and their scale
(|C|(|B|) + |A||C| + |B||A| with 1-norms of the sheared vertices) is 1.4375 + O(y0),
so with the relative margin 1e-5 the ray grazes the edge y = 0 exactly when
|y0| < 1.4375e-5: offsets of +-1e-6 graze, +-1e-4 do not, and 1.3e-5 / 1.6e-5 bracket
the threshold within 10% (a dropped product term in the scale moves it by 13-65%)."""
import numpy as np
import pytest

from goodphotons_amd import abi, scenes
from tests.parity import GRAZE_MARGIN, classify_prim_mismatches

INF = np.inf
T0 = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]


def _scene(tris):
    b = scenes._Builder()
    for t in tris:
        b.tri(*t, 0)
    v, m = b.arrays()
    spd = abi.spectrum(abi.SPEC_BLACKBODY, 6500.0, 1.0, band=(380.0, 780.0))
    em = scenes.quad_emitter((10, 10, 10), (0.1, 0, 0), (0, 0, 0.1), (0, -1, 0), spd)
    mats = [scenes.diffuse(abi.spectrum(abi.SPEC_CONST, 0.5))]
    return scenes.SceneData("margin", v, m, mats, [em], scenes._std_camera(8))


def _down(xy):
    xy = np.atleast_2d(np.asarray(xy, np.float64))
    return np.concatenate([xy, np.ones((len(xy), 1)), np.tile([0.0, 0.0, -1.0], (len(xy), 1))], 1)


def _report(orc, tris, xy, query):
    cs = orc.scene(_scene(tris))
    return orc.closest_hit_ambiguity(cs, _down(xy), np.asarray(query, np.int32), rel_margin=GRAZE_MARGIN)


@pytest.mark.parametrize("y0,prim,t,edge,t_robust,tq,q_graze", [
    (1e-6, 0, 1.0, True, INF, 1.0, True),        # inside, grazing the edge
    (-1e-6, -1, INF, False, INF, 1.0, True),     # outside, grazing: a miss, grown hit at 1
    (1e-4, 0, 1.0, False, 1.0, 1.0, False),      # robustly inside
    (-1e-4, -1, INF, False, INF, INF, False),    # robustly outside
    (1.3e-5, 0, 1.0, True, INF, 1.0, True),      # just inside the margin
    (1.6e-5, 0, 1.0, False, 1.0, 1.0, False),    # just outside the margin
])
def test_silhouette_offsets(orc, y0, prim, t, edge, t_robust, tq, q_graze):
    r = _report(orc, [T0], [(0.25, y0)], [0])
    assert r["prim"][0] == prim and r["t"][0] == t and r["edge"][0] == edge
    # t2 is the nearest grown hit of a triangle other than the winner: with one
    # triangle it exists only when that triangle is not the winner but grown-hit
    assert r["t2"][0] == (1.0 if (prim < 0 and np.isfinite(tq)) else INF)
    assert r["t_robust"][0] == t_robust and r["tq"][0] == tq and r["q_graze"][0] == q_graze


def test_exact_tie_duplicate_triangle(orc):
    # two identical triangles: the first in input order wins (strict t < best), the
    # other is an exact tie at the same distance, nothing grazes
    r = _report(orc, [T0, T0], [(0.25, 0.25)], [1])
    assert r["prim"][0] == 0 and r["t"][0] == 1.0 and r["t2"][0] == 1.0
    assert not r["edge"][0] and r["t_robust"][0] == 1.0 and r["tq"][0] == 1.0 and not r["q_graze"][0]


T1 = [(1, 0, 0), (1, 1, 0), (0, 1, 0)]      # shares the diagonal x + y = 1 with T0


def test_ray_through_shared_coplanar_edge(orc):
    # exactly on the shared diagonal: both are hit (the watertight test keeps edges),
    # the first wins, the other ties at the same distance, and the winner is grazed
    r = _report(orc, [T0, T1], [(0.5, 0.5), (0.5 + 1e-6, 0.5), (0.6, 0.6)], [1, 0, 0])
    assert list(r["prim"]) == [0, 1, 1] and list(r["t"]) == [1.0, 1.0, 1.0]
    assert list(r["t2"]) == [1.0, 1.0, INF]
    assert list(r["edge"]) == [True, True, False]
    assert list(r["tq"]) == [1.0, 1.0, INF] and list(r["q_graze"]) == [True, True, False]


@pytest.mark.parametrize("delta,tie", [(1e-7, True), (1e-5, False)])
def test_ray_near_folded_edge(orc, delta, tie):
    # T0 in z = 0 and a fold rising from its edge y = 0 along the plane z = -y: the ray
    # through (0.25, -delta) hits the fold at t = 1 - delta (inside it) and T0's grown
    # copy at t = 1: a near-tie only when delta is within 1e-6 relative of t
    fold = [(0, 0, 0), (1, 0, 0), (0, -1, 1)]
    r = _report(orc, [T0, fold], [(0.25, -delta)], [0])
    assert r["prim"][0] == 1 and r["t"][0] == pytest.approx(1.0 - delta, abs=1e-15)
    assert r["t2"][0] == 1.0 and r["tq"][0] == 1.0 and r["q_graze"][0]
    assert (abs(r["t2"][0] - r["t"][0]) <= 1e-6 * r["t"][0]) == tie


WALL = [[(-5, -5, -1), (5, -5, -1), (5, 5, -1)], [(-5, -5, -1), (5, 5, -1), (-5, 5, -1)]]


def _classify(orc, tris, xy, gprim, gt):
    sd = _scene(tris)
    cs = orc.scene(sd)
    rays = _down(xy)
    cprim, ct = cs.closest_hit(rays)
    counts, labels = classify_prim_mismatches(orc, cs, rays.astype(np.float32), np.asarray(gprim, np.int32),
                                              np.asarray(gt, np.float32), cprim, ct)
    return cprim, list(labels)


def test_classification_of_planted_answers(orc):
    """Answers planted as the 'GPU' side: valid answers on either side of a graze are
    labelled edge / miss_graze, a tie near_tie, anything else unexplained."""
    wall_a = 2   # the wall triangle (y < x) containing (0.25, +-1e-6) and (0.25, 0.1) at z = -1
    tris = [T0, T1] + WALL
    cprim, lab = _classify(orc, tris,
                           [(0.25, 1e-6), (0.25, -1e-6), (0.25, 0.1), (0.5, 0.5), (0.25, 0.1)],
                           [wall_a, 0, wall_a, 1, 1], [2.0, 1.0, 2.0, 1.0, 1.0])
    assert list(cprim) == [0, wall_a, 0, 0, 0]
    assert lab == ["edge", "edge", "unexplained", "near_tie", "unexplained"]


def test_classification_with_misses(orc):
    cprim, lab = _classify(orc, [T0], [(0.25, -1e-6), (0.25, 1e-6), (0.25, 0.25), (0.25, -0.25)],
                           [0, -1, -1, 0], [1.0, INF, INF, 1.0])
    assert list(cprim) == [-1, 0, 0, -1]
    assert lab == ["miss_graze", "miss_graze", "unexplained", "unexplained"]


def test_classification_checks_the_gpu_distance(orc):
    # the right tied triangle at a wrong distance is not a tie
    cprim, lab = _classify(orc, [T0, T0], [(0.25, 0.25)], [1], [1.01])
    assert lab == ["unexplained"]
