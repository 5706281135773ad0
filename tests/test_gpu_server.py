"""Continuous-batching server: each job's film and ledger are the same alone and in
the mix (bit-identical: per-job photon streams + integer film accumulation), jobs
admitted while others are in flight finish, admission respects the HBM budget."""
import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import scenes

pytestmark = pytest.mark.gpu


def _jobs():
    # (scenes.mirror_box(32), 1 << 15, 8, 0, 55) joins this mix once the flat-mirror
    # connection is switched back on
    return [
        (scenes.cornell_c0(64), 1 << 16, 8, 3, 11),
        (scenes.caustics_c1(48), 1 << 16, 16, 0, 22),
        (scenes.glossy_box(32), 1 << 15, 8, 0, 33),
        (scenes.cornell_c0(16), 3000, 8, 0, 44),
    ]


def test_alone_equals_in_the_mix(gplib):
    specs = _jobs()
    gscenes = [gp.GpuScene(sd, gplib) for sd, *_ in specs]
    alone = []
    for (sd, n, depth, rr, seed), gs in zip(specs, gscenes):
        srv = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=8, lib=gplib)
        jid = srv.submit(gs, sd.camera, n, seed_base=seed, max_depth=depth, rr_depth=rr)
        srv.step()
        alone.append(srv.take(jid, sd.camera.res_x, sd.camera.res_y))
        srv.close()
    srv = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=8, lib=gplib)
    ids = [srv.submit(gs, sd.camera, n, seed_base=seed, max_depth=depth, rr_depth=rr)
           for (sd, n, depth, rr, seed), gs in zip(specs, gscenes)]
    srv.step()
    mixed = [srv.take(j, sd.camera.res_x, sd.camera.res_y) for j, (sd, *_) in zip(ids, specs)]
    srv.close()
    for (ra, xa, ha), (rm, xm, hm) in zip(alone, mixed):
        assert ra.status == 0 and rm.status == 0
        assert np.array_equal(xa, xm) and np.array_equal(ha, hm)
        for k in ("emitted", "absorbed", "escaped", "residual"):
            assert getattr(ra.energy, k) == getattr(rm.energy, k)
        assert ra.stats.segments == rm.stats.segments


def test_server_job_equals_gp_trace(gplib):
    sd = scenes.caustics_c1(48)
    gs = gp.GpuScene(sd, gplib)
    n = 1 << 16
    x1, h1, e1, s1 = gp.render_forward(gs, sd.camera, n, seed_base=5, hero_c=4, max_depth=16)
    srv = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=4, lib=gplib)
    jid = srv.submit(gs, sd.camera, n, seed_base=5, max_depth=16)
    srv.step()
    r, x2, h2 = srv.take(jid, 48, 48)
    srv.close()
    assert np.array_equal(x1, x2) and np.array_equal(h1, h2)
    assert r.stats.segments == s1["segments"]


def test_late_arrivals_join_running_wavefront(gplib):
    sd = scenes.cornell_c0(32)
    gs = gp.GpuScene(sd, gplib)
    srv = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=4, lib=gplib)
    # larger than the path queue, so it is still emitting after the first graph chunk
    first = srv.submit(gs, sd.camera, 1 << 23, seed_base=0, max_depth=8)
    srv.step(max_iterations=4)          # first job now in flight
    q, live = srv.pending()
    assert live == 1
    late = [srv.submit(gs, sd.camera, 1 << 14, seed_base=1000 * k, max_depth=8) for k in range(1, 6)]
    srv.step()
    assert srv.pending() == (0, 0)
    assert sorted(srv.poll()) == sorted([first] + late)
    res = [srv.take(j, 32, 32, film=False)[0] for j in [first] + late]
    assert srv.poll() == []
    assert all(r.status == 0 for r in res)
    assert all(r.stats.photons == n for r, n in zip(res, [1 << 23] + [1 << 14] * 5))
    assert all(r.finish_ms >= r.start_ms >= r.submit_ms for r in res)
    srv.close()


def test_admission_by_hbm_budget(gplib):
    sd = scenes.cornell_c0(64)
    gs = gp.GpuScene(sd, gplib)
    film_bytes = 64 * 64 * 32
    srv = gp.Server(hbm_budget_bytes=film_bytes + 8192, max_live_jobs=8, lib=gplib)
    # depth 16: a job cannot retire within the first 8-iteration chunk
    ids = [srv.submit(gs, sd.camera, 1 << 14, seed_base=k << 14, max_depth=16) for k in range(3)]
    srv.step(max_iterations=2)
    q, live = srv.pending()
    assert live == 1 and q == 2              # only one job's film fits at a time
    srv.step()
    assert srv.pending() == (0, 0)
    for j in ids:
        r, _, _ = srv.take(j, 64, 64, film=False)
        assert r.status == 0
    big = scenes.cornell_c0(256)
    with pytest.raises(gp.GpError) as ei:
        srv.submit(gs, big.camera, 100, max_depth=8)
    assert ei.value.status == 7
    srv.close()


def test_engine_is_exclusive_while_a_server_has_jobs_in_flight(gplib):
    # One control block per device: while a server holds live jobs, a trace, a tuning
    # change or a second server is refused (GP_ERR_STATE) instead of clobbering them;
    # once the server has drained, traces work again and the server's result is unchanged.
    sd = scenes.cornell_c0(32)
    gs = gp.GpuScene(sd, gplib)
    ref = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=4, lib=gplib)
    rj = ref.submit(gs, sd.camera, 1 << 22, seed_base=3, max_depth=8)
    ref.step()
    r0, x0, h0 = ref.take(rj, 32, 32)
    ref.close()
    srv = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=4, lib=gplib)
    jid = srv.submit(gs, sd.camera, 1 << 22, seed_base=3, max_depth=8)
    srv.step(max_iterations=4)
    assert srv.pending()[1] == 1
    f = gp.Film(32, 32, max_photons=1 << 12, lib=gplib)
    with pytest.raises(gp.GpError) as ei:
        gp.trace(gs, sd.camera, f, 1 << 12, max_depth=8)
    assert ei.value.status == 8
    with pytest.raises(gp.GpError) as ei:
        gplib.set_tuning(iters_per_graph=8)
    assert ei.value.status == 8
    other = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=4, lib=gplib)
    other.submit(gs, sd.camera, 1 << 12, seed_base=9, max_depth=8)
    with pytest.raises(gp.GpError) as ei:
        other.step()
    assert ei.value.status == 8
    srv.step()
    r1, x1, h1 = srv.take(jid, 32, 32)
    assert np.array_equal(x0, x1) and np.array_equal(h0, h1)
    gp.trace(gs, sd.camera, f, 1 << 12, max_depth=8)      # the device is free again
    other.step()
    other.close()
    f.close()
    srv.close()


def test_server_resumes_after_a_trace_used_the_engine(gplib):
    # A trace between two of a server's jobs resets the shared control block; the server
    # notices (engine epoch) and starts its next job from a fresh one.
    sd = scenes.cornell_c0(32)
    gs = gp.GpuScene(sd, gplib)
    alone = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=4, lib=gplib)
    aj = alone.submit(gs, sd.camera, 1 << 16, seed_base=77, max_depth=8)
    alone.step()
    ra, xa, ha = alone.take(aj, 32, 32)
    alone.close()
    srv = gp.Server(hbm_budget_bytes=1 << 30, max_live_jobs=4, lib=gplib)
    j1 = srv.submit(gs, sd.camera, 1 << 16, seed_base=5, max_depth=8)
    srv.step()
    srv.take(j1, 32, 32)
    gp.render_forward(gs, sd.camera, 1 << 16, seed_base=11, max_depth=8)
    j2 = srv.submit(gs, sd.camera, 1 << 16, seed_base=77, max_depth=8)
    srv.step()
    rb, xb, hb = srv.take(j2, 32, 32)
    srv.close()
    assert np.array_equal(xa, xb) and np.array_equal(ha, hb)
