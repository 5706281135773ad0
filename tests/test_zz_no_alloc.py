"""gp_film_tonemap_noalloc (the tone map without device allocations: its working memory
is a fixed scratch array in the library) against gp_film_tonemap, both called through
the C ABI: bit-identical 8-bit images, exposures and anchors, on two calls in a row
(the second on a smaller, non-square film, while the scratch array still holds the
first call's keys and histogram state), over every exposure mode (p99 auto exposure,
exposure compensation, absolute gain, locked anchor); a film beyond the scratch limit
(2048 x 2048 pixels) is refused with GP_ERR_CAPACITY."""
import ctypes as C

import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import scenes

pytestmark = pytest.mark.gpu

GP_ERR_CAPACITY = 6


def _traced_film(gplib, rx, ry, n):
    sd = scenes.cornell_c0(rx).with_camera_res(rx, ry)
    gs = gp.GpuScene(sd, gplib)
    f = gp.Film(rx, ry, max_photons=n, lib=gplib)
    gp.trace(gs, sd.camera, f, n, hero_c=4, max_depth=8, rr_depth=3)
    gs.close()
    return f


def _call(L, name, film, n, comp, absolute, anchor):
    img = np.full((film.ry, film.rx, 3), 7, np.uint8)
    ex = C.c_double(-1.0)
    anc = C.c_double(anchor)
    st = getattr(L.lib, name)(film.h, n, comp, absolute, C.byref(anc), img.ctypes.data_as(C.POINTER(C.c_uint8)),
                              C.byref(ex))
    return st, img, ex.value, anc.value


MODES = [(0.0, 0, 0.0), (2.0, 0, 0.0), (0.0, 1, 0.0), (1.5, 0, 0.37)]   # (comp, absolute, anchor)


def test_noalloc_tonemap_equals_tonemap(gplib):
    L = gplib
    for rx, ry, n in ((512, 512, 1 << 20), (96, 40, 1 << 16)):
        f = _traced_film(gplib, rx, ry, n)
        print(f"[noalloc] film {rx}x{ry}, {n} photons", flush=True)
        for comp, absolute, anchor in MODES:
            sa, ia, ea, aa = _call(L, "gp_film_tonemap", f, n, comp, absolute, anchor)
            sb, ib, eb, ab = _call(L, "gp_film_tonemap_noalloc", f, n, comp, absolute, anchor)
            assert sa == 0 and sb == 0
            assert np.array_equal(ia, ib), (rx, ry, comp, absolute, anchor)
            assert ea == eb and aa == ab
            assert ia.any()
        f.close()


def test_noalloc_tonemap_refuses_films_beyond_its_scratch(gplib):
    f = gp.Film(2049, 2048, max_photons=16, lib=gplib)
    st, _, _, _ = _call(gplib, "gp_film_tonemap_noalloc", f, 16, 0.0, 0, 0.0)
    f.close()
    assert st == GP_ERR_CAPACITY
