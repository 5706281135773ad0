"""Ray binning (gp_tuning.ray_sort: 0 off, 1 on, 2 automatic): reordering each
iteration's paths by origin cell and direction octant changes which lanes trace which
paths, never what a path computes,
and the film, ledger and counters are integer sums. So every result must be
bit-identical with and without it: films (fixed-point words), ledgers, segment /
connection / splat counts and the shadow rays' traversal statistics (the closest-hit
node visits and triangle tests depend on which rays share a wave, so they only agree
to within 10%), for single traces (host and
device trees, flat mirrors, several cameras), small queues that wrap many times, and
the server's job mix, and a server whose automatic binning switches on and off while
paths are in flight."""
import contextlib

import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import scenes

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def tuning(lib, **kw):
    saved = lib.tuning()
    lib.set_tuning(**kw)
    try:
        yield
    finally:
        lib.set_tuning(**{k: getattr(saved, k) for k, _ in saved._fields_})


def _trace(lib, gs, sd, n, depth, rr=0, stats=False, sort=0, **kw):
    with tuning(lib, ray_sort=sort, **kw):
        f = gp.Film(sd.camera.res_x, sd.camera.res_y, max_photons=n, lib=lib)
        e = gp.abi.Energy()
        _, st = gp.trace(gs, sd.camera, f, n, hero_c=4, max_depth=depth, rr_depth=rr, energy=e,
                         traversal_stats=stats)
        words = f.raw()[0].copy()
        f.close()
    return words, (e.emitted, e.absorbed, e.sensor, e.escaped, e.residual), st


def _same(a, b):
    wa, ea, sa = a
    wb, eb, sb = b
    assert np.array_equal(wa, wb)
    assert ea == eb
    for k in ("photons", "segments", "connections", "splats", "direct_connections", "shadow_node_visits",
              "shadow_tri_tests"):
        assert getattr(sa, k) == getattr(sb, k), k
    # the closest-hit counts depend on which rays share a wave (postponed leaves)
    for k in ("node_visits", "tri_tests"):
        assert abs(getattr(sa, k) - getattr(sb, k)) <= 0.1 * max(1, getattr(sa, k)), k


CASES = {
    "c0_rr3": (lambda: scenes.cornell_c0(64), 1 << 18, 8, 3),
    "c1": (lambda: scenes.caustics_c1(64), 1 << 18, 16, 0),
    "glossy": (lambda: scenes.glossy_box(64), 1 << 17, 8, 0),
    "mirror": (lambda: scenes.mirror_box(64), 1 << 17, 8, 0),
    "c2_l5": (lambda: scenes.icosphere_c2(5, 64), 1 << 18, 12, 0),
    "soup20k": (lambda: scenes.soup(20000, 5, 64), 1 << 17, 8, 0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_binned_trace_is_bit_identical(gplib, name):
    make, n, depth, rr = CASES[name]
    sd = make()
    gs = gp.GpuScene(sd, gplib)
    try:
        a = _trace(gplib, gs, sd, n, depth, rr, sort=0)
        b = _trace(gplib, gs, sd, n, depth, rr, sort=1)
        print(f"{name}: segments {a[2].segments} connections {a[2].connections}", flush=True)
        assert a[2].segments > n
        _same(a, b)
    finally:
        gs.close()


def test_binned_traversal_statistics_agree(gplib):
    sd = scenes.caustics_c1(64)
    gs = gp.GpuScene(sd, gplib)
    try:
        a = _trace(gplib, gs, sd, 1 << 17, 16, stats=True, sort=0)
        b = _trace(gplib, gs, sd, 1 << 17, 16, stats=True, sort=1)
        assert a[2].node_visits > 0 and a[2].shadow_tri_tests > 0
        _same(a, b)
    finally:
        gs.close()


def test_binned_small_queue_and_grid(gplib):
    # a 2^12-path queue refilled many times, on a 4-waves-per-CU grid
    sd = scenes.caustics_c1(32)
    gs = gp.GpuScene(sd, gplib)
    try:
        a = _trace(gplib, gs, sd, 1 << 16, 16, sort=0, queue_capacity=4096, grid_waves_per_cu=4)
        b = _trace(gplib, gs, sd, 1 << 16, 16, sort=1, queue_capacity=4096, grid_waves_per_cu=4)
        assert a[2].iterations > 16
        _same(a, b)
        c = _trace(gplib, gs, sd, 1 << 16, 16, sort=1)
        _same(a, c)
    finally:
        gs.close()


def test_binned_device_tree(gplib):
    sd = scenes.caustics_c1(64)
    gs = gp.GpuScene(sd, gplib, builder="device")
    try:
        _same(_trace(gplib, gs, sd, 1 << 17, 16, sort=0), _trace(gplib, gs, sd, 1 << 17, 16, sort=1))
    finally:
        gs.close()


def test_binned_shared_cameras(gplib):
    sd = scenes.cornell_c0(48)
    cams = [sd.camera, scenes.camera((0.2, 0.6, 2.5), (0.5, 0.4, 0.5), (0, 1, 0), 45.0, 40, 32),
            scenes.camera((0.9, 0.9, 1.9), (0.4, 0.3, 0.4), (0, 1, 0), 30.0, 24, 24)]
    gs = gp.GpuScene(sd, gplib)
    try:
        out = []
        for sort in (0, 1):
            with tuning(gplib, ray_sort=sort):
                films, e, st = gp.render_forward_shared(gs, cams, 1 << 16, hero_c=4, max_depth=8)
            out.append(([x for x, _ in films], [h for _, h in films], st["segments"], st["splats"]))
        for k in range(3):
            assert np.array_equal(out[0][0][k], out[1][0][k]) and np.array_equal(out[0][1][k], out[1][1][k])
        assert out[0][2:] == out[1][2:]
    finally:
        gs.close()


def test_binned_server_mix(gplib):
    specs = [(scenes.cornell_c0(48), 1 << 16, 8, 3, 11), (scenes.caustics_c1(48), 1 << 16, 16, 0, 22),
             (scenes.mirror_box(32), 1 << 15, 8, 0, 55), (scenes.icosphere_c2(4, 32), 1 << 15, 12, 0, 66)]
    gscenes = [gp.GpuScene(sd, gplib) for sd, *_ in specs]
    try:
        res = []
        for sort in (0, 1):
            with tuning(gplib, ray_sort=sort):
                srv = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=8, lib=gplib)
                ids = [srv.submit(gs, sd.camera, n, seed_base=seed, max_depth=depth, rr_depth=rr)
                       for (sd, n, depth, rr, seed), gs in zip(specs, gscenes)]
                srv.step()
                res.append([srv.take(j, sd.camera.res_x, sd.camera.res_y) for j, (sd, *_) in zip(ids, specs)])
                srv.close()
        for (ra, xa, ha), (rb, xb, hb) in zip(*res):
            assert np.array_equal(xa, xb) and np.array_equal(ha, hb)
            assert ra.stats.segments == rb.stats.segments and ra.stats.splats == rb.stats.splats
    finally:
        for gs in gscenes:
            gs.close()


def test_automatic_binning_switches_under_a_running_server(gplib):
    # ray_sort = 2 bins while a live job's scene has 4096 triangles or more: a server
    # that starts with a small scene (not binned) and then admits a large one switches
    # binning on with paths of the first job still in the queue (their keys were never
    # written), and off again when the large job ends while small ones still run (a
    # 2^14-path queue spreads the jobs over many iterations); every film equals the
    # unbinned server's.
    small = (scenes.cornell_c0(48), 1 << 17, 8, 3, 11)
    large = (scenes.icosphere_c2(4, 32), 1 << 15, 12, 0, 66)
    tail = (scenes.cornell_c0(32), 1 << 17, 8, 0, 77)
    assert large[0].n_tris >= 4096 > small[0].n_tris
    specs = [small, large, tail]
    gscenes = [gp.GpuScene(sd, gplib) for sd, *_ in specs]
    try:
        res = []
        for sort in (0, 2):
            with tuning(gplib, ray_sort=sort, queue_capacity=1 << 14):
                srv = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=8, lib=gplib)
                ids = [srv.submit(gscenes[0], small[0].camera, small[1], seed_base=small[4], max_depth=small[2],
                                  rr_depth=small[3])]
                srv.step(max_iterations=4)
                for (sd, n, depth, rr, seed), gs in zip(specs[1:], gscenes[1:]):
                    ids.append(srv.submit(gs, sd.camera, n, seed_base=seed, max_depth=depth, rr_depth=rr))
                srv.step()
                res.append([srv.take(j, sd.camera.res_x, sd.camera.res_y) for j, (sd, *_) in zip(ids, specs)])
                srv.close()
        for (ra, xa, ha), (rb, xb, hb) in zip(*res):
            assert np.array_equal(xa, xb) and np.array_equal(ha, hb)
            assert ra.stats.segments == rb.stats.segments and ra.stats.splats == rb.stats.splats
    finally:
        for gs in gscenes:
            gs.close()
