"""Camera connections through flat mirrors on the device (k_mirror) against the
oracle's restatement of the reference's connectSpecularPlane (src/render.h:1334-1376;
the oracle matches the reference's own mirror-scene films bit for bit). The pixels
where the camera sees a mirror receive nothing but mirror connections (a delta
surface has no direct connection), so comparing those pixels isolates the connector
from the rest of the image."""
import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import scenes

pytestmark = pytest.mark.gpu

MIRROR = 4   # material index of the mirrors in scenes.mirror_box


def _mirror_pixels(orc, sd):
    """Pixels whose four corner rays from the eye all hit a mirror triangle first."""
    cam = sd.camera
    rx, ry = cam.res_x, cam.res_y
    b = orc.camera_basis(cam)
    eye, u, v, w, thx, thy = b[0:3], b[3:6], b[6:9], b[9:12], b[12], b[13]
    xs = np.arange(rx + 1) / rx * 2 - 1
    ys = np.arange(ry + 1) / ry * 2 - 1
    X, Y = np.meshgrid(xs, ys)
    d = w + np.outer(X.ravel() * thx, u) + np.outer(Y.ravel() * thy, v)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([np.broadcast_to(eye, d.shape), d], 1)
    prim, _ = orc.scene(sd).closest_hit(rays)
    hit = (prim >= 0) & (sd.mat_ids[np.maximum(prim, 0)] == MIRROR)
    hit = hit.reshape(ry + 1, rx + 1)
    return hit[:-1, :-1] & hit[1:, :-1] & hit[:-1, 1:] & hit[1:, 1:]


@pytest.mark.parametrize("hero_c", [4, 1])
def test_mirror_pixels_match_oracle(gplib, orc, hero_c):
    sd = scenes.mirror_box(64)
    n = 1 << 19
    print(f"[mirror] hero_c={hero_c} n={n}", flush=True)
    gs = gp.GpuScene(sd, gplib)
    gx, gh, ge, gst = gp.render_forward(gs, sd.camera, n, hero_c=hero_c, max_depth=8)
    cx, ch, ce, cst = orc.render(orc.scene(sd), sd.camera, n, hero_c=hero_c, max_bounce=8, threads=8)
    m = _mirror_pixels(orc, sd).ravel()
    assert m.sum() > 300
    G, R = gx[m], cx[m]
    rel = [np.sqrt(np.sum((G[:, c] - R[:, c]) ** 2) / np.sum(R[:, c] ** 2)) for c in range(3)]
    lum = G[:, 1].sum() / R[:, 1].sum()
    share = R[:, 1].sum() / cx[:, 1].sum()
    print(f"[mirror] {int(m.sum())} mirror pixels, {share:.3f} of the light, rel {rel}, lum {lum:.5f}",
          flush=True)
    assert share > 0.03
    assert max(rel) < 0.01 and abs(lum - 1) < 0.003
    assert abs(gh[m].sum() / ch[m].sum() - 1) < 2e-3
    assert abs(gst["connections"] / cst["connections"] - 1) < 2e-3


def test_mirror_film_bit_reproducible_and_progressive(gplib):
    sd = scenes.mirror_box(48)
    gs = gp.GpuScene(sd, gplib)
    n = 1 << 17
    words = []
    for _ in range(2):
        f = gp.Film(48, 48, max_photons=n, lib=gplib)
        gp.trace(gs, sd.camera, f, n, hero_c=4, max_depth=8)
        words.append(f.raw()[0])
        f.close()
    assert np.array_equal(words[0], words[1])
    f = gp.Film(48, 48, max_photons=n, lib=gplib)
    gp.trace(gs, sd.camera, f, n // 2, seed_base=0, hero_c=4, max_depth=8)
    gp.trace(gs, sd.camera, f, n // 2, seed_base=n // 2, hero_c=4, max_depth=8)
    assert np.array_equal(words[0], f.raw()[0])
    f.close()


def test_scene_without_mirrors_unchanged_after_mirror_scene(gplib):
    # The mirror-vertex queue stays allocated once a mirror scene ran; a scene without
    # mirrors must trace exactly as before (no mirror kernel, same film words).
    sd = scenes.cornell_c0(32)
    gs = gp.GpuScene(sd, gplib)
    a = gp.Film(32, 32, max_photons=1 << 16, lib=gplib)
    gp.trace(gs, sd.camera, a, 1 << 16, hero_c=4, max_depth=8)
    ms = scenes.mirror_box(32)
    gm = gp.GpuScene(ms, gplib)
    f = gp.Film(32, 32, max_photons=1 << 15, lib=gplib)
    gp.trace(gm, ms.camera, f, 1 << 15, hero_c=4, max_depth=8)
    b = gp.Film(32, 32, max_photons=1 << 16, lib=gplib)
    gp.trace(gs, sd.camera, b, 1 << 16, hero_c=4, max_depth=8)
    assert np.array_equal(a.raw()[0], b.raw()[0])
