"""Write the golden vectors that pin the oracle, from the reference's own code.

Runs the reference renderer's forward path compiled from its headers
(oracle/_ref/libref.so, built by oracle/Makefile where the reference sources
exist) on seeded inputs and stores inputs + outputs as small .npz fixtures next
to this script. The oracle tests compare the repository's restatement with these
bit for bit. Re-run with:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from goodphotons_amd import abi, scenes  # noqa: E402


def seeded_rays(sd, n, seed):
    """Rays from points inside the scene's bounding box (float32-exact) in uniform
    directions, the -checkbvh pattern (src/main.cpp:448-473) with a local rng."""
    rng = np.random.default_rng(seed)
    v = sd.verts.reshape(-1, 3).astype(np.float64)
    lo, hi = v.min(0), v.max(0)
    o = lo + (hi - lo) * rng.uniform(-0.1, 1.1, size=(n, 3))
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    rays = np.concatenate([o, d], 1).astype(np.float32).astype(np.float64)
    return rays


def main():
    R = oracle.Reference()
    out = {}
    # --- rng: PCG32 streams and per-photon seeding
    seqs = [(0, 0), (54, 42), (0xDEADBEEF, 0x12345678), (2**63 + 5, 2**62 + 7)]
    out["pcg_seq"] = np.array([s for s, _ in seqs], np.uint64)
    out["pcg_seed"] = np.array([t for _, t in seqs], np.uint64)
    out["pcg_out"] = np.stack([R.pcg_stream(s, t, 32) for s, t in seqs])
    ks = [0, 1, 2, 1000, 2**31, 2**40 + 3]
    out["unit_k"] = np.array(ks, np.uint64)
    out["unit_out"] = np.stack([R.seed_unit_uniforms(k, abi.REFERENCE_SALT, 16) for k in ks])
    xs = [0, 1, 0x9E3779B97F4A7C15, 2**64 - 1]
    out["mix_in"] = np.array(xs, np.uint64)
    out["mix_out"] = np.array([R.mix64(x) for x in xs], np.uint64)
    # --- spectra
    lam = np.linspace(360.0, 830.0, 941)
    out["lam"] = lam
    out["cie"] = R.cie(lam)
    out["cie_y_integral"] = np.array([R.cie_y_integral()])
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
        out["spec_" + k] = R.spectrum_eval(s, lam)
    # --- emission sampler + hero bundle
    u = (np.arange(4096) + 0.5) / 4096.0
    u = np.concatenate([u, [0.0, 1e-9, 0.25, 0.5, 0.75, 1 - 2**-24]])
    l, p, integral = R.emission_sample(specs["bb6500band"], u)
    out["emit_u"], out["emit_lam"], out["emit_pdf"] = u, l, p
    out["emit_integral"] = np.array([integral])
    bl, bp = [], []
    for uu in u[::37]:
        ok, la, pd = R.sample_bundle(specs["bb6500band"], uu, 4)
        bl.append(la if ok else np.full(4, np.nan))
        bp.append(pd if ok else np.full(4, np.nan))
    out["bundle_u"], out["bundle_lam"], out["bundle_pdf"] = u[::37], np.array(bl), np.array(bp)
    # --- sampling lobes
    out["cos_n"] = np.array([0.0, -1.0, 0.0])
    out["cos_dirs"] = R.cosine_hemisphere(out["cos_n"], 11, 5, 64)
    md = np.array([0.3, 0.8, -0.52])
    md /= np.linalg.norm(md)
    out["glossy_m"] = md
    out["glossy_dirs"] = R.sample_glossy(md, 0.3, 13, 9, 64)
    # --- camera
    cam = c0.camera
    out["cam_basis"] = R.camera_basis(cam)
    pts = np.random.default_rng(3).uniform(-0.2, 1.2, size=(256, 3))
    px, py, cc, d2, om = R.camera_project(cam, pts)
    out["cam_pts"], out["cam_px"], out["cam_py"] = pts, px, py
    out["cam_cos"], out["cam_d2"], out["cam_omega"] = cc, d2, om
    np.savez_compressed(os.path.join(HERE, "primitives.npz"), **out)

    # --- closest hit + traversal statistics on seeded ray batches
    for name, sd, n in (("c0", scenes.cornell_c0(64), 4096), ("c1", scenes.caustics_c1(64), 4096)):
        s = R.scene(sd)
        rays = seeded_rays(sd, n, 17)
        prim, t, nv, lt = s.closest_hit(rays, stats=True)
        md = np.full(n, 10.0)
        occ_rays = rays.copy()
        occ = s.occluded(occ_rays, md * 0.5)
        np.savez_compressed(os.path.join(HERE, f"hits_{name}.npz"), rays=rays, prim=prim, t=t,
                            node_visits=nv, leaf_tests=lt, occ_maxdist=md * 0.5, occluded=occ,
                            bvh_nodes=np.array([s.bvh_nodes]), total_power=np.array([s.total_power]))

    # --- films and energy ledgers (32x32, one thread so accumulation order is fixed)
    cases = [("c0", scenes.cornell_c0(32), 4096), ("c1", scenes.caustics_c1(32), 4096),
             ("glossy", scenes.glossy_box(32), 4096), ("mirror", scenes.mirror_box(32), 4096)]
    for name, sd, n in cases:
        s = R.scene(sd)
        for hc in (1, 4):
            xyz, hits, e, rq = R.render(s, sd.camera, n, seed_base=12345, hero_c=hc,
                                        max_bounce=sd.max_depth, threads=1)
            np.savez_compressed(os.path.join(HERE, f"film_{name}_h{hc}.npz"), xyz=xyz, hits=hits,
                                energy=np.array([e[k] for k in ("emitted", "absorbed", "sensor", "escaped", "residual")]),
                                ray_queries=np.array([rq["ray_queries"]]), n=np.array([n]),
                                seed_base=np.array([12345]), max_bounce=np.array([sd.max_depth]))
    # --- flat mirror planes (Scene::mirrorPlanes) of the mirror scene
    sd = scenes.mirror_box(32)
    np.savez_compressed(os.path.join(HERE, "mirror_planes.npz"), planes=R.scene(sd).mirror_planes())
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
