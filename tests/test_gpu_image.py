"""gp_film_to_linear and gp_film_tonemap (the device reduction of filmToLinear /
filmToRgb8, src/main.cpp:10282-10348) against the oracle on the same film words:
bit-exact float32 linear radiance and bit-exact 8-bit pixels, including the p99
anchor found by the device radix select, exposure compensation, the locked anchor and
absolute EV."""
import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import scenes

pytestmark = pytest.mark.gpu


def _traced(sd, res, n, lib):
    s = sd.with_camera_res(res)
    gs = gp.GpuScene(s, lib)
    film = gp.Film(res, res, max_photons=n, lib=lib)
    gp.trace(gs, s.camera, film, n, hero_c=s.hero_c, max_depth=s.max_depth, rr_depth=s.rr_depth)
    return film


@pytest.mark.parametrize("case", ["c0_64", "c1_200x200", "c1_1024"])
def test_film_to_image_matches_oracle(case, orc, gplib):
    if case == "c0_64":
        film, n = _traced(scenes.cornell_c0(64), 64, 1 << 16, gplib), 1 << 16
    elif case == "c1_200x200":
        film, n = _traced(scenes.caustics_c1(200), 200, 1 << 18, gplib), 1 << 18
    else:
        film, n = _traced(scenes.caustics_c1(1024), 1024, 1 << 20, gplib), 1 << 20
    print(f"[image] {case}", flush=True)
    xyz, _ = film.download()
    w, h = film.rx, film.ry
    lin = film.to_linear(n)
    ref_lin = orc.film_to_linear(xyz, n).astype(np.float32).reshape(h, w, 3)
    assert np.array_equal(lin, ref_lin)
    img, exposure, anchor = film.tonemap(n)
    ref_img, ref_e, ref_x, _ = orc.film_to_rgb8(xyz, w, h, n)
    assert anchor == ref_e and exposure == ref_x
    assert np.array_equal(img, ref_img)
    assert img.max() > 0
    # exposure compensation, locked anchor, absolute EV
    img2, x2, a2 = film.tonemap(n, exposure_comp=0.25, anchor=anchor * 3.0)
    ref2, _, rx2, _ = orc.film_to_rgb8(xyz, w, h, n, exp_comp=0.25, anchor=anchor * 3.0)
    assert a2 == anchor * 3.0 and x2 == rx2 and np.array_equal(img2, ref2)
    img3, x3, _ = film.tonemap(n, exposure_comp=2.0, absolute=True)
    ref3, _, rx3, _ = orc.film_to_rgb8(xyz, w, h, n, exp_comp=2.0, absolute=True)
    assert x3 == rx3 == 12.0 and np.array_equal(img3, ref3)
