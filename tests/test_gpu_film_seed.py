"""Session companions of the C ABI: gp_film_download_hits (the camera-0 noise poll,
sharedForwardGpuHits0, src/render_cuda.h:135-137) returns exactly the hit counts of a
full download; gp_film_seed (sharedForwardGpuBegin's seedFilms, resuming from running
totals held as doubles, e.g. a checkpoint the reference wrote) puts the sums back into
the same fixed-point words when the scene, camera and photon budget are the same, so
a render resumed from its downloaded totals (or from an FTBUF01 file without the
fixed-point section) continues bit-exactly; invalid seeds are refused and leave the
film untouched."""
import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import scenes

pytestmark = pytest.mark.gpu


CASES = {
    "c0_rr3": (lambda: scenes.cornell_c0(64), 1 << 16, 8, 3),
    "c1": (lambda: scenes.caustics_c1(64), 1 << 16, 16, 0),
    "mirror": (lambda: scenes.mirror_box(48), 1 << 15, 8, 0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_seeded_film_continues_bit_exactly(gplib, name):
    make, n, depth, rr = CASES[name]
    sd = make()
    gs = gp.GpuScene(sd, gplib)
    budget = 4 * n
    try:
        full = gp.Film(sd.camera.res_x, sd.camera.res_y, max_photons=budget, lib=gplib)
        gp.trace(gs, sd.camera, full, n, seed_base=0, hero_c=4, max_depth=depth, rr_depth=rr)
        gp.trace(gs, sd.camera, full, n, seed_base=n, hero_c=4, max_depth=depth, rr_depth=rr)
        half = gp.Film(sd.camera.res_x, sd.camera.res_y, max_photons=budget, lib=gplib)
        gp.trace(gs, sd.camera, half, n, seed_base=0, hero_c=4, max_depth=depth, rr_depth=rr)
        x, h = half.download()
        assert np.array_equal(half.download_hits(), h)
        seeded = gp.Film(sd.camera.res_x, sd.camera.res_y, max_photons=budget, lib=gplib)
        seeded.seed(gs, sd.camera, n, x, h, max_depth=depth)
        assert np.array_equal(seeded.raw()[0], half.raw()[0])
        gp.trace(gs, sd.camera, seeded, n, seed_base=n, hero_c=4, max_depth=depth, rr_depth=rr)
        assert np.array_equal(seeded.raw()[0], full.raw()[0])
        assert np.array_equal(seeded.download_hits(), full.download()[1])
        for f in (full, half, seeded):
            f.close()
    finally:
        gs.close()


def test_seed_from_checkpoint_without_fixed_point_section(gplib, orc, tmp_path):
    # A file holding only the FTBUF01 doubles, as the reference writes it: the core's
    # loader refuses to resume it directly; its doubles (read as the reference's
    # readCheckpoint does) seed a film that then continues exactly.
    sd = scenes.cornell_c0(48)
    gs = gp.GpuScene(sd, gplib)
    n, depth = 1 << 16, 8
    try:
        a = gp.Film(48, 48, max_photons=2 * n, lib=gplib)
        e = gp.abi.Energy()
        gp.trace(gs, sd.camera, a, n, hero_c=4, max_depth=depth, energy=e)
        path = str(tmp_path / "ref.ftbuf")
        a.save_checkpoint(path, n, e, guard=7)
        raw = open(path, "rb").read()
        cut = raw.index(b"GPFILM01")
        open(path, "wb").write(raw[:cut])
        with pytest.raises(gp.GpError):
            gp.Film.load_checkpoint(path, guard=7, lib=gplib)
        rc, n_ck, en, xyz, hits, trailing = orc.read_checkpoint(path, 48, 48, 7)
        assert rc == 0 and n_ck == n and trailing == 0 and en[0] == e.emitted
        b = gp.Film(48, 48, max_photons=2 * n, lib=gplib)
        b.seed(gs, sd.camera, n_ck, xyz, hits, max_depth=depth)
        assert np.array_equal(b.raw()[0], a.raw()[0])
        gp.trace(gs, sd.camera, a, n, seed_base=n, hero_c=4, max_depth=depth)
        gp.trace(gs, sd.camera, b, n, seed_base=n, hero_c=4, max_depth=depth)
        assert np.array_equal(b.raw()[0], a.raw()[0])
        a.close()
        b.close()
    finally:
        gs.close()


def test_invalid_seeds_are_refused(gplib):
    sd = scenes.cornell_c0(16)
    gs = gp.GpuScene(sd, gplib)
    npx = 16 * 16
    try:
        f = gp.Film(16, 16, max_photons=1 << 16, lib=gplib)
        ok_x, ok_h = np.zeros((npx, 3)), np.zeros(npx)
        for x, h in ((np.full((npx, 3), -1.0), ok_h), (np.full((npx, 3), np.nan), ok_h),
                     (np.full((npx, 3), np.inf), ok_h), (ok_x, np.full(npx, 0.5)), (ok_x, np.full(npx, -1.0)),
                     (np.full((npx, 3), 1e300), ok_h)):
            with pytest.raises(gp.GpError):
                f.seed(gs, sd.camera, 1000, x, h, max_depth=8)
            assert not f.raw()[0].any() and f.raw()[1] == 0.0
        gp.trace(gs, sd.camera, f, 1000, hero_c=4, max_depth=8)
        with pytest.raises(gp.GpError):   # already traced into
            f.seed(gs, sd.camera, 1000, ok_x, ok_h, max_depth=8)
        wrong = scenes.cornell_c0(32).camera
        g = gp.Film(16, 16, max_photons=1 << 16, lib=gplib)
        with pytest.raises(gp.GpError):   # camera resolution differs from the film's
            g.seed(gs, wrong, 1000, ok_x, ok_h, max_depth=8)
        f.close()
        g.close()
    finally:
        gs.close()
