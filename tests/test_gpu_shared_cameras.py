"""One photon flight splatted into several cameras (gp_trace_shared, the device twin
of renderForwardShared, src/main.cpp:11354-11405): every camera's film is
bit-identical to a single-camera trace of the same photons (the reference states the
same property for its model-B shared render), the ledger and segment count are
those of one flight, connections add up over the cameras, and the path works in the
same engine as single-camera traces (queues re-sized for the camera count)."""
import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import scenes

pytestmark = pytest.mark.gpu


def _cams(sd):
    base = sd.camera
    views = [((0.5, 0.5, 2.7), (0.5, 0.5, 0.5), 48, 48),
             ((1.6, 0.7, 2.2), (0.5, 0.45, 0.5), 64, 40),
             ((-0.4, 0.9, 1.8), (0.5, 0.4, 0.4), 32, 32)]
    out = []
    for eye, tgt, rx, ry in views:
        out.append(scenes.camera(eye, tgt, tuple(base.up), base.fov_y_deg, rx, ry))
    return out


@pytest.mark.parametrize("name,make,n,depth,rr", [
    ("c0", lambda: scenes.cornell_c0(48), 1 << 17, 8, 3),
    ("c1", lambda: scenes.caustics_c1(48), 1 << 17, 16, 0),
    ("mirror", lambda: scenes.mirror_box(48), 1 << 17, 8, 0),
])
def test_shared_cameras_equal_single_camera_traces(gplib, name, make, n, depth, rr):
    sd = make()
    gs = gp.GpuScene(sd, gplib)
    cams = _cams(sd)
    films, e_sh, st_sh = gp.render_forward_shared(gs, cams, n, seed_base=7, hero_c=4, max_depth=depth, rr_depth=rr)
    conns = 0
    for cam, (x, h) in zip(cams, films):
        xs, hs, es, sts = gp.render_forward(gs, cam, n, seed_base=7, hero_c=4, max_depth=depth, rr_depth=rr)
        assert np.array_equal(x, xs) and np.array_equal(h, hs)
        assert sts["segments"] == st_sh["segments"] and sts["photons"] == st_sh["photons"]
        for k in ("emitted", "absorbed", "escaped", "residual"):
            assert es[k] == e_sh[k]
        conns += sts["connections"]
        assert x[:, 1].max() > 0
    assert conns == st_sh["connections"]
    print(f"[shared] {name}: {len(cams)} cameras, segments {st_sh['segments']}, connections {conns}", flush=True)


def test_shared_rejects_bad_arguments(gplib):
    sd = scenes.cornell_c0(16)
    gs = gp.GpuScene(sd, gplib)
    f = gp.Film(16, 16, max_photons=1 << 10, lib=gplib)
    with pytest.raises(gp.GpError):          # one film for two cameras
        gp.trace_shared(gs, [sd.camera, sd.camera], [f, f], 1 << 10, max_depth=8)
    with pytest.raises(gp.GpError):          # more than 16 cameras
        films = [gp.Film(16, 16, max_photons=1 << 10, lib=gplib) for _ in range(17)]
        gp.trace_shared(gs, [sd.camera] * 17, films, 1 << 10, max_depth=8)


def test_refused_shared_trace_leaves_every_film_unchanged(gplib):
    """A job whose second film refuses its photons (budget used up) must not charge the
    first film: that film still takes its full budget afterwards."""
    sd = scenes.cornell_c0(16)
    gs = gp.GpuScene(sd, gplib)
    cam2 = _cams(sd)[1]
    n = 1 << 12
    fa = gp.Film(16, 16, max_photons=n, lib=gplib)
    fb = gp.Film(cam2.res_x, cam2.res_y, max_photons=n, lib=gplib)
    gp.trace(gs, cam2, fb, n, max_depth=8)                     # fb's budget is now spent
    with pytest.raises(gp.GpError) as err:
        gp.trace_shared(gs, [sd.camera, cam2], [fa, fb], n, max_depth=8)
    assert gp.abi.STATUS_NAMES[err.value.status] == "GP_ERR_CAPACITY"
    gp.trace(gs, sd.camera, fa, n, max_depth=8)                # fa still holds n photons
