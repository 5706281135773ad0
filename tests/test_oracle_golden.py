"""The oracle (oracle/gp_oracle.cpp) against the reference's golden vectors and
closed forms. All comparisons with golden vectors are bit-exact."""
import numpy as np
import pytest

from goodphotons_amd import abi, scenes
from tests.conftest import golden


def test_pcg32_published_stream(orc):
    # O'Neill's pcg32 demo: srandom(initstate=42, initseq=54) -> these six outputs.
    out = orc.pcg_stream(54, 42, 6)
    assert [hex(x) for x in out] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330",
                                     "0x83d2f293", "0xbfa4784b", "0xcbed606e"]


def test_splitmix64_published_value(orc):
    # First output of SplitMix64 seeded with 0 (Steele, Lea, Flood 2014).
    assert orc.mix64(0) == 0xE220A8397B1DCDAF


def test_rng_golden(orc):
    g = golden("primitives.npz")
    for i, (seq, seed) in enumerate(zip(g["pcg_seq"], g["pcg_seed"])):
        assert np.array_equal(orc.pcg_stream(int(seq), int(seed), 32), g["pcg_out"][i])
    for i, k in enumerate(g["unit_k"]):
        assert np.array_equal(orc.seed_unit_uniforms(int(k), abi.REFERENCE_SALT, 16), g["unit_out"][i])
    for x, y in zip(g["mix_in"], g["mix_out"]):
        assert orc.mix64(int(x)) == int(y)


def test_spectra_golden(orc):
    g = golden("primitives.npz")
    lam = g["lam"]
    assert np.array_equal(orc.cie(lam), g["cie"])
    assert orc.cie_y_integral() == g["cie_y_integral"][0]
    c0 = scenes.cornell_c0(32)
    c1 = scenes.caustics_c1(32)
    specs = {
        "white": c0.materials[0].reflect, "red": c0.materials[1].reflect,
        "green": c0.materials[2].reflect, "bk7": c1.materials[4].ior, "sf10": c1.materials[5].ior,
        "bb6500band": c0.emitters[0].spd,
        "cauchy": abi.spectrum(abi.SPEC_CAUCHY, 1.5, 0.004),
        "bb3200": abi.spectrum(abi.SPEC_BLACKBODY, 3200.0, 1.0),
    }
    for k, s in specs.items():
        assert np.array_equal(orc.spectrum_eval(s, lam), g["spec_" + k]), k


def test_glass_catalogue_values(orc):
    # Schott catalogue: N-BK7 nd=1.51680 nF=1.52238 nC=1.51432; SF10 nd=1.72825.
    c1 = scenes.caustics_c1(32)
    bk7 = orc.spectrum_eval(c1.materials[4].ior, np.array([587.5618, 486.1327, 656.2725]))
    assert np.allclose(bk7, [1.51680, 1.52238, 1.51432], atol=2e-5)
    sf10 = orc.spectrum_eval(c1.materials[5].ior, np.array([587.5618]))
    assert abs(sf10[0] - 1.72825) < 1e-4
    cau = orc.spectrum_eval(abi.spectrum(abi.SPEC_CAUCHY, 1.5, 0.004), np.array([500.0]))
    assert abs(cau[0] - (1.5 + 0.004 / 0.25)) < 1e-15


def test_planck_wien_peak(orc):
    lam = np.linspace(380, 780, 40001)
    bb = orc.spectrum_eval(abi.spectrum(abi.SPEC_BLACKBODY, 6500.0, 1.0), lam)
    assert abs(lam[bb.argmax()] - 2.897771955e-3 / 6500 * 1e9) < 0.02


def test_cie_y_integral_close_to_tabulated(orc):
    # The tabulated CIE 1931 ybar sums to ~106.86 over 360..830 at 1 nm; the
    # Wyman-Sloan-Shirley fit is within 0.1%.
    assert abs(orc.cie_y_integral() / 106.857 - 1) < 1e-3


def test_emission_and_bundle_golden(orc):
    g = golden("primitives.npz")
    spd = scenes.cornell_c0(32).emitters[0].spd
    lam, pdf, integral = orc.emission_sample(spd, g["emit_u"])
    assert np.array_equal(lam, g["emit_lam"]) and np.array_equal(pdf, g["emit_pdf"])
    assert integral == g["emit_integral"][0]
    # band-limited SPD: every hero wavelength with positive pdf is inside the band
    assert np.all((lam[pdf > 0] >= 380.0) & (lam[pdf > 0] <= 781.0))
    for u, bl, bp in zip(g["bundle_u"], g["bundle_lam"], g["bundle_pdf"]):
        ok, la, pd = orc.sample_bundle(spd, float(u), 4)
        if np.isnan(bl[0]):
            assert not ok
        else:
            assert ok and np.array_equal(la, bl) and np.array_equal(pd, bp)


def test_sampling_lobes_golden(orc):
    g = golden("primitives.npz")
    assert np.array_equal(orc.cosine_hemisphere(g["cos_n"], 11, 5, 64), g["cos_dirs"])
    assert np.array_equal(orc.sample_glossy(g["glossy_m"], 0.3, 13, 9, 64), g["glossy_dirs"])
    # cosine-weighted: mean cos about n is 2/3 (closed form), all in the hemisphere
    d = orc.cosine_hemisphere(g["cos_n"], 3, 4, 20000)
    c = d @ g["cos_n"]
    assert np.all(c > 0) and abs(c.mean() - 2 / 3) < 0.01


def test_camera_golden(orc):
    g = golden("primitives.npz")
    cam = scenes.cornell_c0(32).camera
    assert np.array_equal(orc.camera_basis(cam), g["cam_basis"])
    px, py, cc, d2, om = orc.camera_project(cam, g["cam_pts"])
    assert np.array_equal(px, g["cam_px"]) and np.array_equal(py, g["cam_py"])
    assert np.array_equal(cc, g["cam_cos"]) and np.array_equal(d2, g["cam_d2"])
    assert np.array_equal(om, g["cam_omega"])


@pytest.mark.parametrize("name,builder", [("c0", scenes.cornell_c0), ("c1", scenes.caustics_c1)])
def test_closest_hit_golden(orc, name, builder):
    g = golden(f"hits_{name}.npz")
    sd = builder(64)
    s = orc.scene(sd)
    assert s.bvh_nodes == g["bvh_nodes"][0]
    assert s.total_power == g["total_power"][0]
    prim, t, nv, lt = s.closest_hit(g["rays"], stats=True)
    assert np.array_equal(prim, g["prim"])
    assert np.array_equal(t, g["t"])
    assert np.array_equal(nv, g["node_visits"]) and np.array_equal(lt, g["leaf_tests"])
    assert np.array_equal(s.occluded(g["rays"], g["occ_maxdist"]), g["occluded"])
    # brute force agrees with the BVH: the reference's -checkbvh invariant (same hit
    # distance within 1e-7 and same material, src/main.cpp:462-464). Prim ids may
    # differ only on exact ties, e.g. a box's bottom face lying on the floor.
    lp, lt2 = s.closest_hit_linear(g["rays"])
    assert np.array_equal(lp >= 0, prim >= 0)
    hit = prim >= 0
    assert np.all(np.abs(lt2[hit] - t[hit]) < 1e-7)
    mats = sd.mat_ids
    assert np.array_equal(mats[lp[hit]], mats[prim[hit]])
    ties = hit & (lp != prim)
    assert np.all(lt2[ties] == t[ties])


def test_axis_aligned_wall_closed_form(orc):
    # A ray straight down from (0.2, 0.5, 0.7) hits the c0 floor at t = 0.5 exactly.
    s = orc.scene(scenes.cornell_c0(32))
    prim, t = s.closest_hit(np.array([[0.2, 0.5, 0.7, 0.0, -1.0, 0.0]]))
    assert prim[0] in (0, 1) and t[0] == 0.5


@pytest.mark.parametrize("name,builder", [("c0", scenes.cornell_c0), ("c1", scenes.caustics_c1),
                                          ("glossy", scenes.glossy_box), ("mirror", scenes.mirror_box)])
@pytest.mark.parametrize("hero", [1, 4])
def test_film_golden(orc, name, builder, hero):
    g = golden(f"film_{name}_h{hero}.npz")
    sd = builder(32)
    s = orc.scene(sd)
    xyz, hits, e, st = orc.render(s, sd.camera, int(g["n"][0]), seed_base=int(g["seed_base"][0]),
                                  hero_c=hero, max_bounce=int(g["max_bounce"][0]), threads=1)
    assert np.array_equal(xyz, g["xyz"]) and np.array_equal(hits, g["hits"])
    ev = np.array([e[k] for k in ("emitted", "absorbed", "sensor", "escaped", "residual")])
    assert np.array_equal(ev, g["energy"])
    assert st["segments"] + st["connections"] == g["ray_queries"][0]


def test_ledger_closes_without_dispersion(orc):
    sd = scenes.cornell_c0(32)
    s = orc.scene(sd)
    _, _, e, _ = orc.render(s, sd.camera, 20000, hero_c=4, max_bounce=8, threads=4)
    total = e["absorbed"] + e["escaped"] + e["residual"] + e["sensor"]
    assert abs(total / e["emitted"] - 1.0) < 1e-12


def test_film_independent_of_thread_count(orc):
    sd = scenes.cornell_c0(32)
    s = orc.scene(sd)
    a = orc.render(s, sd.camera, 8000, hero_c=4, max_bounce=8, threads=1)
    b = orc.render(s, sd.camera, 8000, hero_c=4, max_bounce=8, threads=3)
    assert np.allclose(a[0], b[0], rtol=1e-12, atol=0) and np.array_equal(a[1], b[1])


def test_rr_depth_is_unbiased(orc):
    # Deferring Russian roulette changes variance, not the expectation.
    sd = scenes.cornell_c0(16)
    s = orc.scene(sd)
    n = 200000
    a = orc.render(s, sd.camera, n, hero_c=4, max_bounce=8, rr_depth=0, threads=8)[0].sum(0)
    b = orc.render(s, sd.camera, n, hero_c=4, max_bounce=8, rr_depth=3, seed_base=n, threads=8)[0].sum(0)
    assert np.all(np.abs(b / a - 1) < 0.01)


def test_mirror_planes_golden(orc):
    # Coplanar mirror triangles collapse to one plane each (the short box's six faces
    # and the wall panel), with the reference's canonical normal, offset and padded box.
    g = golden("mirror_planes.npz")["planes"]
    p = orc.scene(scenes.mirror_box(32)).mirror_planes()
    assert len(g) == 7 and np.array_equal(p, g)


def test_mirror_plane_geometry_closed_form(orc):
    # Closed forms: a unit normal with its dominant component positive, every member
    # vertex on the plane (|n.v - d| at float rounding), and the wall panel at x = 0.001.
    sd = scenes.mirror_box(32)
    p = orc.scene(sd).mirror_planes()
    n, d = p[:, :3], p[:, 3]
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-15)
    lead = n[np.arange(len(n)), np.argmax(np.abs(n), axis=1)]
    assert np.all(lead > 0)
    v = sd.verts.reshape(-1, 3, 3).astype(np.float64)[sd.mat_ids == 4]
    for tri in v:
        k = np.argmin([np.abs(tri @ nn - dd).max() for nn, dd in zip(n, d)])
        assert np.abs(tri @ n[k] - d[k]).max() < 1e-6
        assert np.all(tri >= p[k, 4:7]) and np.all(tri <= p[k, 7:10])
    wall = p[np.argmax(np.abs(n[:, 0]) * (np.abs(d - 0.001) < 1e-6))]
    assert np.allclose(wall[:3], [1, 0, 0]) and abs(wall[3] - 0.001) < 1e-9


def test_mirror_connection_adds_light(orc):
    # The mirror panel shows the lit room: the same scene with the mirror material
    # replaced by a black diffuse one has strictly less light, and the difference
    # lands in the panel's columns of the image (the left wall).
    sd = scenes.mirror_box(32)
    a = orc.render(orc.scene(sd), sd.camera, 40000, hero_c=4, max_bounce=8, threads=8)[0]
    mats = list(sd.materials)
    mats[4] = scenes.diffuse(abi.spectrum(abi.SPEC_CONST, 0.0))
    sd.materials = mats
    b = orc.render(orc.scene(sd), sd.camera, 40000, hero_c=4, max_bounce=8, threads=8)[0]
    ya, yb = a[:, 1].reshape(32, 32), b[:, 1].reshape(32, 32)
    assert ya.sum() > 1.1 * yb.sum()
    assert ya[:, :8].sum() > 2.0 * yb[:, :8].sum()
