"""GPU (HIP, gfx950) against the oracle on the same inputs, through the C ABI.

Closest-hit probe: primitive ids bit-exact except on rays whose two nearest hits lie
within 1e-6 relative distance (near-ties); rays where the float32 test decides a
boundary graze differently from the double one are counted apart, each checked to be
the correct answer on one side of the graze, and bounded (tests/parity.py
check_prim_parity); any other mismatch fails. Films: the GPU draws the
same photons as the CPU (same per-photon PCG streams), so at equal photon count
the per-channel relative RMSE must be < 1% (observed ~1e-4, float32 path
divergence only) and mean luminance within 0.3%.
"""
import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import abi, scenes
from tests.parity import check_prim_parity, image_metrics, seeded_rays, to_sampled

pytestmark = pytest.mark.gpu

PROBE_CASES = [
    ("c0", lambda: scenes.cornell_c0(64), 200000),
    ("c1", lambda: scenes.caustics_c1(64), 100000),
    ("c2_l5", lambda: scenes.icosphere_c2(level=5, res=64), 50000),
    ("soup20k", lambda: scenes.soup(20000, 7, res=64), 50000),
]


@pytest.mark.parametrize("name,builder,n", PROBE_CASES, ids=[c[0] for c in PROBE_CASES])
def test_closest_hit_prim_ids(gplib, orc, name, builder, n):
    sd = builder()
    print(f"[probe] {name} tris={sd.n_tris} rays={n}", flush=True)
    _check_closest(gplib, orc, sd, seeded_rays(sd, n, 1234), name)


def _axis_rays(sd, n, seed):
    """Rays along the six axis directions exactly (inverse direction components are
    +-inf), a fifth of them starting exactly on a plane of the scene's bounds."""
    rng = np.random.default_rng(seed)
    v = sd.verts.reshape(-1, 3).astype(np.float32)
    lo, hi = v.min(0), v.max(0)
    o = (lo + (hi - lo) * rng.uniform(0.0, 1.0, size=(n, 3))).astype(np.float32)
    snap = rng.uniform(size=n) < 0.2
    ax = rng.integers(0, 3, n)
    side = rng.integers(0, 2, n)
    o[snap, ax[snap]] = np.where(side[snap] == 0, lo[ax[snap]], hi[ax[snap]])
    d = np.zeros((n, 3), np.float32)
    da = rng.integers(0, 3, n)
    d[np.arange(n), da] = np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0)
    return np.concatenate([o, d], 1).astype(np.float32)


@pytest.mark.parametrize("name,builder", [("c0", lambda: scenes.cornell_c0(64)),
                                          ("c1", lambda: scenes.caustics_c1(64))], ids=["c0", "c1"])
def test_axis_aligned_rays_match_oracle(gplib, orc, name, builder):
    """Axis-parallel rays make the slab test divide by zero (infinite inverse
    components); hits must still agree with the oracle's prim ids."""
    sd = builder()
    _check_closest(gplib, orc, sd, _axis_rays(sd, 20000, 77), name + "_axis", max_frac=0.2)


def _check_closest(gplib, orc, sd, rays, name, max_frac=0.1):
    gs = gp.GpuScene(sd, gplib)
    cs = orc.scene(sd)
    gprim, gt = gs.closest_hit(rays, tmin=1e-6)
    cprim, ct = cs.closest_hit(rays.astype(np.float64), tmin=1e-6)
    check_prim_parity(orc, cs, rays, gprim, gt, cprim, ct, name, max_tie_frac=max_frac)


@pytest.mark.parametrize("axis", [False, True], ids=["uniform_dirs", "axis_dirs"])
def test_any_hit_matches_oracle(gplib, orc, axis):
    sd = scenes.caustics_c1(64)
    gs = gp.GpuScene(sd, gplib)
    cs = orc.scene(sd)
    rays = _axis_rays(sd, 50000, 99) if axis else seeded_rays(sd, 50000, 99)
    rng = np.random.default_rng(5)
    tmax = rng.uniform(0.05, 1.5, len(rays)).astype(np.float32)
    occ_g = gs.any_hit(rays, tmax, tmin=1e-6)
    # the oracle's occluded() tests [tmin, maxDist - tmin); pass maxDist = tmax + tmin
    occ_c = cs.occluded(rays.astype(np.float64), tmax.astype(np.float64) + 1e-6)
    _, ct = cs.closest_hit(rays.astype(np.float64))
    near = np.abs(ct - tmax) <= 2e-6 * (1.0 + tmax)
    bad = (occ_g != occ_c) & ~near
    assert bad.sum() == 0, f"{int(bad.sum())} occlusion mismatches away from the segment end"


FILM_CASES = [
    ("c0", lambda: scenes.cornell_c0(64), 1 << 18, 8, 3),
    ("c0_rr0", lambda: scenes.cornell_c0(64), 1 << 18, 8, 0),
    ("c1", lambda: scenes.caustics_c1(64), 1 << 18, 16, 0),
    ("glossy", lambda: scenes.glossy_box(64), 1 << 17, 8, 0),
    ("mirror", lambda: scenes.mirror_box(64), 1 << 18, 8, 0),
]


@pytest.mark.parametrize("name,builder,n,depth,rr", FILM_CASES,
                         ids=["c0", "c0_rr0", "c1", "glossy", "mirror"])
def test_film_matches_oracle(gplib, orc, name, builder, n, depth, rr):
    sd = builder()
    print(f"[film] {name} n={n} depth={depth} rr={rr}", flush=True)
    gs = gp.GpuScene(sd, gplib)
    gx, gh, ge, gst = gp.render_forward(gs, sd.camera, n, hero_c=4, max_depth=depth, rr_depth=rr)
    cs = orc.scene(sd)
    cx, ch, ce, cst = orc.render(cs, sd.camera, n, hero_c=4, max_bounce=depth, rr_depth=rr, threads=8)
    m = image_metrics(gx, cx, n, n, res=(sd.camera.res_x, sd.camera.res_y))
    print(f"[film] {name}: {m} gpu_segments={gst['segments']} cpu_segments={cst['segments']}", flush=True)
    assert m["rel_rmse_max"] < 0.01
    assert abs(m["mean_lum_ratio"] - 1) < 0.003
    # same photons: emitted power identical, path statistics within float noise
    assert abs(ge["emitted"] / ce["emitted"] - 1) < 1e-6
    assert abs(gst["segments"] / cst["segments"] - 1) < 2e-3
    assert abs(gst["connections"] / cst["connections"] - 1) < 2e-3
    assert abs(gh.sum() / ch.sum() - 1) < 2e-3
    for k in ("absorbed", "escaped", "residual"):
        assert abs(ge[k] - ce[k]) <= 2e-3 * ce["emitted"], k


@pytest.mark.parametrize("name,make,n,depth", [("c0", lambda: scenes.cornell_c0(64), 1 << 18, 8),
                                               ("c1", lambda: scenes.caustics_c1(64), 1 << 18, 16)],
                         ids=["c0", "c1"])
def test_sampled_spectra_film_matches_oracle(gplib, orc, name, make, n, depth):
    """Every spectrum as 470 samples (SPEC_SAMPLED, the form the reference's Spectrum
    functions take through the adapter): the device interpolates float samples, the
    oracle double ones; same photons, same bar as the analytic spectra. The sampled
    scene also stays close to the analytic one (interpolation between 1-nm samples)."""
    sd = make()
    ss = to_sampled(sd, orc)
    gs = gp.GpuScene(ss, gplib)
    gx, gh, ge, gst = gp.render_forward(gs, ss.camera, n, hero_c=4, max_depth=depth, rr_depth=sd.rr_depth)
    cx, ch, ce, cst = orc.render(orc.scene(ss), ss.camera, n, hero_c=4, max_bounce=depth, rr_depth=sd.rr_depth,
                                 threads=8)
    ax, ah, ae, ast = orc.render(orc.scene(sd), sd.camera, n, hero_c=4, max_bounce=depth, rr_depth=sd.rr_depth,
                                 threads=8)
    m = image_metrics(gx, cx, n, n, res=(64, 64))
    a = image_metrics(cx, ax, n, n, res=(64, 64))
    print(f"[film] {name} sampled: gpu vs oracle {m}; sampled vs analytic (oracle) {a}", flush=True)
    assert m["rel_rmse_max"] < 0.01 and abs(m["mean_lum_ratio"] - 1) < 0.003
    assert abs(ge["emitted"] / ce["emitted"] - 1) < 1e-6
    assert abs(gst["segments"] / cst["segments"] - 1) < 2e-3
    assert a["rel_rmse_max"] < 0.01 and abs(a["mean_lum_ratio"] - 1) < 0.003


def test_energy_ledger_closes(gplib):
    sd = scenes.cornell_c0(64)
    gs = gp.GpuScene(sd, gplib)
    _, _, e, _ = gp.render_forward(gs, sd.camera, 1 << 18, hero_c=4, max_depth=8, rr_depth=0)
    total = e["absorbed"] + e["escaped"] + e["residual"] + e["sensor"]
    assert abs(total / e["emitted"] - 1.0) < 1e-6


def test_film_bit_reproducible(gplib):
    sd = scenes.caustics_c1(64)
    gs = gp.GpuScene(sd, gplib)
    words = []
    for _ in range(2):
        f = gp.Film(64, 64, max_photons=1 << 18, lib=gplib)
        gp.trace(gs, sd.camera, f, 1 << 18, hero_c=4, max_depth=16)
        words.append(f.raw()[0])
        f.close()
    assert np.array_equal(words[0], words[1])


def test_progressive_batches_equal_single_shot(gplib):
    # seed_base = photons already traced: two half batches are the same photons as
    # one full batch, and integer accumulation makes the films bit-identical.
    sd = scenes.cornell_c0(64)
    gs = gp.GpuScene(sd, gplib)
    n = 1 << 17
    a = gp.Film(64, 64, max_photons=n, lib=gplib)
    gp.trace(gs, sd.camera, a, n, hero_c=4, max_depth=8)
    b = gp.Film(64, 64, max_photons=n, lib=gplib)
    gp.trace(gs, sd.camera, b, n // 2, seed_base=0, hero_c=4, max_depth=8)
    gp.trace(gs, sd.camera, b, n // 2, seed_base=n // 2, hero_c=4, max_depth=8)
    assert np.array_equal(a.raw()[0], b.raw()[0])


def test_queue_capacity_does_not_change_result(gplib):
    sd = scenes.cornell_c0(32)
    gs = gp.GpuScene(sd, gplib)
    old = gplib.tuning()
    try:
        out = []
        for cap in (1 << 14, 1 << 16):
            gplib.set_tuning(queue_capacity=cap)
            f = gp.Film(32, 32, max_photons=1 << 17, lib=gplib)
            gp.trace(gs, sd.camera, f, 1 << 17, hero_c=4, max_depth=8)
            out.append(f.raw()[0])
        assert np.array_equal(out[0], out[1])
    finally:
        gplib.set_tuning(**{k: getattr(old, k) for k, _ in old._fields_})


@pytest.mark.parametrize("name,make,n,depth", [("c1", lambda: scenes.caustics_c1(32), 1 << 16, 16),
                                               ("mirror", lambda: scenes.mirror_box(32), 1 << 15, 8),
                                               ("c2_l5", lambda: scenes.icosphere_c2(5, 32), 1 << 16, 12)],
                         ids=["c1", "mirror", "c2_l5"])
def test_lds_stack_depths_give_identical_films(gplib, name, make, n, depth):
    """Every LDS stack depth (8, 12, 16, 24, 32 entries; the closest-hit, any-hit, mirror
    and probe kernels each have their own variants and spill beyond them) traces the
    same paths: the stack only holds the pending nodes, so films, counters and probe
    prim ids are bit-identical, in queue order and binned."""
    sd = make()
    gs = gp.GpuScene(sd, gplib)
    old = gplib.tuning()
    rays = seeded_rays(sd, 4096, 3)
    try:
        ref = None
        for sort in (0, 1):
            for lds in (8, 12, 16, 24, 32):
                gplib.set_tuning(lds_stack=lds, ray_sort=sort)
                f = gp.Film(32, 32, max_photons=n, lib=gplib)
                _, st = gp.trace(gs, sd.camera, f, n, hero_c=4, max_depth=depth)
                p, t = gs.closest_hit(rays, tmin=1e-6)
                got = (f.raw()[0].copy(), st.segments, st.connections, st.splats, p, t)
                f.close()
                if ref is None:
                    ref = got
                    assert got[1] > n
                else:
                    assert np.array_equal(got[0], ref[0]), (sort, lds)
                    assert got[1:4] == ref[1:4], (sort, lds)
                    assert np.array_equal(got[4], ref[4]) and np.array_equal(got[5], ref[5]), (sort, lds)
    finally:
        gplib.set_tuning(**{k: getattr(old, k) for k, _ in old._fields_})
        gs.close()


@pytest.mark.parametrize("mtype,rough", [(3, 0.0), (abi.MAT_GLOSSY, 0.0), (7, 0.0)],
                         ids=["halfmirror", "glossy_rough0", "grating"])
def test_unsupported_material_rejected(gplib, mtype, rough):
    sd = scenes.cornell_c0(32)
    m = abi.Material()
    m.type = mtype
    m.roughness = rough
    m.reflect = abi.spectrum(abi.SPEC_CONST, 0.9)
    sd.materials[0] = m
    with pytest.raises(gp.GpError) as ei:
        gp.GpuScene(sd, gplib)
    assert ei.value.status == 2


@pytest.mark.parametrize("hero_c", [1, 2, 3])
def test_film_matches_oracle_hero_counts(gplib, orc, hero_c):
    # hero bundles of 1..3 wavelengths (4 is covered above): stratified secondaries,
    # de-hero at the dielectrics and the max-over-lambda roulette all change with C
    sd = scenes.caustics_c1(48)
    n = 1 << 18
    gs = gp.GpuScene(sd, gplib)
    gx, gh, ge, gst = gp.render_forward(gs, sd.camera, n, hero_c=hero_c, max_depth=16)
    cs = orc.scene(sd)
    cx, ch, ce, cst = orc.render(cs, sd.camera, n, hero_c=hero_c, max_bounce=16, threads=8)
    m = image_metrics(gx, cx, n, n, res=(48, 48))
    print(f"[film] c1 hero_c={hero_c}: {m}", flush=True)
    assert m["rel_rmse_max"] < 0.01 and abs(m["mean_lum_ratio"] - 1) < 0.003
    assert abs(ge["emitted"] / ce["emitted"] - 1) < 1e-6
    assert abs(gst["segments"] / cst["segments"] - 1) < 2e-3
    assert abs(gh.sum() / ch.sum() - 1) < 2e-3


def test_non_square_film_matches_oracle(gplib, orc):
    sd = scenes.cornell_c0(64).with_camera_res(80, 36)
    n = 1 << 18
    gs = gp.GpuScene(sd, gplib)
    gx, gh, ge, gst = gp.render_forward(gs, sd.camera, n, hero_c=4, max_depth=8, rr_depth=3)
    cs = orc.scene(sd)
    cx, ch, ce, cst = orc.render(cs, sd.camera, n, hero_c=4, max_bounce=8, rr_depth=3, threads=8)
    m = image_metrics(gx, cx, n, n, res=(80, 36))
    print(f"[film] c0 80x36: {m}", flush=True)
    assert m["rel_rmse_max"] < 0.01 and abs(m["mean_lum_ratio"] - 1) < 0.003


def test_empty_and_tiny_traces(gplib):
    sd = scenes.cornell_c0(16)
    gs = gp.GpuScene(sd, gplib)
    f = gp.Film(16, 16, max_photons=8, lib=gplib)
    e, st = gp.trace(gs, sd.camera, f, 0, max_depth=8)
    assert st.photons == 0 and st.segments == 0 and e.emitted == 0.0
    assert not f.raw()[0].any()
    e, st = gp.trace(gs, sd.camera, f, 1, max_depth=8)
    assert st.photons == 1 and st.segments >= 1 and e.emitted > 0
    e, st = gp.trace(gs, sd.camera, f, 7, seed_base=1, max_depth=8)
    assert st.photons == 7
    with pytest.raises(gp.GpError):           # beyond the film's photon budget
        gp.trace(gs, sd.camera, f, 1, seed_base=8, max_depth=8)


def test_exact_ties_resolve_independently_of_wave_mates(gplib):
    """Every triangle of a soup twice, at identical vertices: every hit is an exact tie.
    The closest-hit traversal postpones leaves until its wave holds one in every lane,
    so a ray's visit order depends on its wave-mates; the answer must not. The same rays
    probed in another order and in a smaller batch return the same primitive and
    distance, and that is the soup's own closest hit (the single-copy scene's primitive,
    modulo the copy, at the same distance: the triangle test does not depend on the
    tree)."""
    import dataclasses
    base = scenes.soup(4000, 11, res=32)
    n = base.n_tris
    sd = dataclasses.replace(base, verts=np.concatenate([base.verts, base.verts]),
                             mat_ids=np.concatenate([base.mat_ids, base.mat_ids]))
    gs = gp.GpuScene(sd, gplib)
    try:
        rays = seeded_rays(sd, 40000, 99)
        p0, t0 = gs.closest_hit(rays, tmin=1e-6)
        perm = np.random.default_rng(5).permutation(len(rays))
        p1, t1 = gs.closest_hit(np.ascontiguousarray(rays[perm]), tmin=1e-6)
        inv = np.empty_like(perm)
        inv[perm] = np.arange(len(perm))
        p1, t1 = p1[inv], t1[inv]
        p2, t2 = gs.closest_hit(np.ascontiguousarray(rays[:3001]), tmin=1e-6)
        hit = p0 >= 0
        print(f"[ties] rays={len(rays)} hits={int(hit.sum())}", flush=True)
        assert hit.sum() > 1000
        assert np.array_equal(p0, p1) and np.array_equal(t0, t1)
        assert np.array_equal(p0[:3001], p2) and np.array_equal(t0[:3001], t2)
    finally:
        gs.close()
    gb = gp.GpuScene(base, gplib)
    try:
        pb, tb = gb.closest_hit(rays, tmin=1e-6)
    finally:
        gb.close()
    assert np.array_equal(pb >= 0, hit)
    assert np.array_equal(p0[hit] % n, pb[hit]) and np.array_equal(t0[hit], tb[hit])
