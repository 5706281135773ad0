"""Traversal statistics of a trace (GP_TRACE_TRAVERSAL_STATS; the reference's
TraversalStats / -bvhstats, src/bvh.h:79, src/main.cpp:2599), on the device.

The counting kernel variants must leave the film and the path statistics unchanged bit
for bit; every query visits at least the root node. The shadow rays' any-hit counts are
a function of the traced rays only, so they repeat exactly across queue capacities; the
closest-hit counts also depend on which rays share a wave (a lane that reached a leaf
keeps visiting nodes while its wave-mates look for one), so with another queue capacity
they repeat within a few percent. Against
the oracle's counts over the reference's binary tree on the same photons: a four-wide
node visit replaces up to three binary inner-node visits and a leaf costs no visit of
its own, so the four-wide tree visits fewer nodes; both trees stop at leaves of at most
four triangles, so the triangle tests stay within a small factor (this library's SAH
over three axes and 32 bins tests 0.4-0.5 times as many as the reference's 16-bin
widest-axis tree on the box scenes)."""
import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import scenes

pytestmark = pytest.mark.gpu

CASES = [("c0", lambda: scenes.cornell_c0(64), 1 << 17, 8),
         ("c1", lambda: scenes.caustics_c1(64), 1 << 17, 16)]


def _run(gs, sd, n, depth, stats):
    f = gp.Film(sd.camera.res_x, sd.camera.res_y, max_photons=n, lib=gs.L)
    try:
        e, st = gp.trace(gs, sd.camera, f, n, hero_c=4, max_depth=depth, rr_depth=sd.rr_depth,
                         traversal_stats=stats)
        return f.raw()[0], st.as_dict()
    finally:
        f.close()


@pytest.mark.parametrize("name,make,n,depth", CASES, ids=[c[0] for c in CASES])
def test_traversal_stats(gplib, orc, name, make, n, depth):
    sd = make()
    gs = gp.GpuScene(sd, gplib)
    w0, s0 = _run(gs, sd, n, depth, False)
    w1, s1 = _run(gs, sd, n, depth, True)
    keys = ("node_visits", "tri_tests", "shadow_node_visits", "shadow_tri_tests")
    print(f"[trav] {name}: " + " ".join(f"{k}={s1[k]}" for k in ("segments", "connections",
                                                                  "direct_connections") + keys), flush=True)
    assert np.array_equal(w0, w1)
    for k in ("photons", "segments", "connections", "splats", "direct_connections"):
        assert s0[k] == s1[k], k
    assert all(s0[k] == 0 for k in keys)
    assert s1["node_visits"] >= s1["segments"] and s1["shadow_node_visits"] >= s1["connections"]
    assert 0 < s1["direct_connections"] <= s1["photons"] < s1["connections"]
    # another queue capacity (other scheduling, other wave-mates): the any-hit counts
    # repeat exactly, the closest-hit counts within 5%
    old = gplib.tuning()
    try:
        gplib.set_tuning(queue_capacity=1 << 15)
        _, s2 = _run(gs, sd, n, depth, True)
    finally:
        gplib.set_tuning(**{k: getattr(old, k) for k, _ in old._fields_})
    print(f"[trav] {name} at 2^15 slots: node_visits={s2['node_visits']} tri_tests={s2['tri_tests']}", flush=True)
    assert s2["shadow_node_visits"] == s1["shadow_node_visits"] and s2["shadow_tri_tests"] == s1["shadow_tri_tests"]
    for k in ("node_visits", "tri_tests"):
        assert abs(s2[k] - s1[k]) <= 0.05 * s1[k], k
    # against the reference's binary tree on the same photons (the oracle's counters)
    _, _, _, c = orc.render(orc.scene(sd), sd.camera, n, hero_c=4, max_bounce=depth, rr_depth=sd.rr_depth,
                            threads=8)
    print(f"[trav] {name} oracle (binary tree): node_visits={c['node_visits']} leaf_tests={c['leaf_tests']} "
          f"any_node_visits={c['any_node_visits']} any_leaf_tests={c['any_leaf_tests']}", flush=True)
    assert s1["segments"] <= s1["node_visits"] < c["node_visits"]
    assert 0.2 < s1["tri_tests"] / c["leaf_tests"] < 1.5
    assert s1["connections"] <= s1["shadow_node_visits"] < c["any_node_visits"]
