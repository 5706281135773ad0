"""Test infrastructure only: ctypes wrappers for the CPU checkers.

  Oracle()      this repository's double-precision restatement (liboracle.so)
  Reference()   the reference renderer's own forward path compiled from its
                headers (oracle/_ref/libref.so; built only where the reference
                sources exist, travels as a built .so)
  RefAdapter()  the drop-in adapter compiled against the reference's headers, on
                scenes built with the reference's own API, beside the reference's
                CPU path (oracle/_ref/libref_adapter.so, built the same way)

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
package, and only as the checker / baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from goodphotons_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


def build(with_ref: bool = True) -> None:
    targets = ["all", "adapter"] if with_ref else [os.path.join(HERE, "liboracle.so")]
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile"), *targets], check=True)


def _ptr(a, t):
    return a.ctypes.data_as(t)


class _CpuRenderer:
    prefix = ""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run oracle build)")
        self.lib = C.CDLL(path)
        p = self.prefix
        L = self.lib
        f = getattr(L, p + "scene_create"); f.restype = C.c_void_p; f.argtypes = [C.POINTER(abi.SceneDesc)]
        f = getattr(L, p + "scene_destroy"); f.restype = None; f.argtypes = [C.c_void_p]
        f = getattr(L, p + "scene_bvh_nodes"); f.restype = C.c_int64; f.argtypes = [C.c_void_p]
        f = getattr(L, p + "scene_total_power"); f.restype = C.c_double; f.argtypes = [C.c_void_p]
        f = getattr(L, p + "scene_mirror_planes"); f.restype = C.c_int
        f.argtypes = [C.c_void_p, _dp, C.c_int]
        f = getattr(L, p + "closest_hit")
        f.argtypes = [C.c_void_p, _dp, C.c_int64, C.c_double, _i32p, _dp, _i64p, _i64p]
        f = getattr(L, p + "closest_hit_linear")
        f.argtypes = [C.c_void_p, _dp, C.c_int64, C.c_double, _i32p, _dp]
        f = getattr(L, p + "occluded"); f.argtypes = [C.c_void_p, _dp, _dp, C.c_int64, _u8p]
        f = getattr(L, p + "pcg_stream"); f.restype = None
        f.argtypes = [C.c_uint64, C.c_uint64, C.c_int, _u32p]
        f = getattr(L, p + "seed_unit_uniforms"); f.restype = None
        f.argtypes = [C.c_uint64, C.c_uint64, C.c_int, _dp]
        f = getattr(L, p + "mix64"); f.restype = C.c_uint64; f.argtypes = [C.c_uint64]
        f = getattr(L, p + "spectrum_eval"); f.restype = None
        f.argtypes = [C.POINTER(abi.Spectrum), _dp, _dp, C.c_int, _dp]
        f = getattr(L, p + "cie"); f.restype = None; f.argtypes = [_dp, C.c_int, _dp]
        f = getattr(L, p + "cie_y_integral"); f.restype = C.c_double; f.argtypes = []
        f = getattr(L, p + "emission_sample"); f.restype = None
        f.argtypes = [C.POINTER(abi.Spectrum), _dp, _dp, C.c_int, _dp, _dp, _dp]
        f = getattr(L, p + "sample_bundle"); f.restype = C.c_int
        f.argtypes = [C.POINTER(abi.Spectrum), _dp, C.c_double, C.c_int, _dp, _dp]
        f = getattr(L, p + "cosine_hemisphere"); f.restype = None
        f.argtypes = [_dp, C.c_uint64, C.c_uint64, C.c_int, _dp]
        f = getattr(L, p + "sample_glossy"); f.restype = None
        f.argtypes = [_dp, C.c_double, C.c_uint64, C.c_uint64, C.c_int, _dp]
        f = getattr(L, p + "camera_basis"); f.restype = None
        f.argtypes = [C.POINTER(abi.CameraDesc), _dp]
        f = getattr(L, p + "camera_project"); f.restype = C.c_int
        f.argtypes = [C.POINTER(abi.CameraDesc), _dp, C.c_int, _i32p, _i32p, _dp, _dp, _dp]
        if p == "or_":   # film -> image: restated from main.cpp (no reference entry point)
            f = L.or_film_to_linear; f.restype = None
            f.argtypes = [_dp, C.c_int64, C.c_double, _dp]
            f = L.or_checkpoint_guard; f.restype = C.c_uint64
            f.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_int, C.c_int, C.c_int]
            f = L.or_read_checkpoint; f.restype = C.c_int
            f.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_uint64, C.c_int, _i64p, _dp, _dp, _dp, _i64p]
            f = L.or_film_to_rgb8; f.restype = None
            f.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _u8p, _dp, _dp]

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    # ---- scenes
    def scene(self, sd) -> "CpuScene":
        return CpuScene(self, sd)

    # ---- primitives
    def pcg_stream(self, seq, seed, n):
        out = np.zeros(n, np.uint32)
        self._f("pcg_stream")(seq, seed, n, _ptr(out, _u32p))
        return out

    def seed_unit_uniforms(self, k, salt, n):
        out = np.zeros(n, np.float64)
        self._f("seed_unit_uniforms")(k, salt, n, _ptr(out, _dp))
        return out

    def mix64(self, x):
        return int(self._f("mix64")(x))

    def spectrum_eval(self, spec, lam, pool=None):
        """pool: the samples of a GP_SPEC_SAMPLED spectrum's table (spec.table indexes it)."""
        lam = np.ascontiguousarray(lam, np.float64)
        out = np.zeros_like(lam)
        pl = None if pool is None else np.ascontiguousarray(pool, np.float64)
        self._f("spectrum_eval")(C.byref(spec), None if pl is None else _ptr(pl, _dp), _ptr(lam, _dp), len(lam),
                                 _ptr(out, _dp))
        return out

    def cie(self, lam):
        lam = np.ascontiguousarray(lam, np.float64)
        out = np.zeros((len(lam), 3))
        self._f("cie")(_ptr(lam, _dp), len(lam), _ptr(out, _dp))
        return out

    def cie_y_integral(self):
        return float(self._f("cie_y_integral")())

    # ---- .ftbuf checkpoints (oracle only)
    def checkpoint_guard(self, n_tris, n_spheres, n_emitters, total_power, mode, rx, ry):
        return int(self.lib.or_checkpoint_guard(n_tris, n_spheres, n_emitters, float(total_power),
                                                ord(mode) if isinstance(mode, str) else int(mode), rx, ry))

    def read_checkpoint(self, path, rx, ry, guard, mode="B"):
        """readCheckpoint: (status, N, energy[5], xyz (rx*ry, 3), hits, trailing bytes)."""
        import os as _os
        n = C.c_int64(0)
        tr = C.c_int64(0)
        en = np.zeros(5)
        xyz = np.zeros((rx * ry, 3))
        hits = np.zeros(rx * ry)
        rc = self.lib.or_read_checkpoint(_os.fsencode(str(path)), rx, ry, int(guard),
                                         ord(mode) if isinstance(mode, str) else int(mode), C.byref(n),
                                         _ptr(en, _dp), _ptr(xyz, _dp), _ptr(hits, _dp), C.byref(tr))
        return rc, n.value, en, xyz, hits, tr.value

    # ---- film -> image (oracle only)
    def film_to_linear(self, xyz, n_photons):
        """xyz: (npix, 3) film sums in row order -> (npix, 3) linear sRGB (double)."""
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        out = np.zeros_like(xyz)
        self.lib.or_film_to_linear(_ptr(xyz, _dp), len(xyz), float(n_photons), _ptr(out, _dp))
        return out

    def film_to_rgb8(self, xyz, w, h, n_photons, exp_comp=0.0, absolute=False, anchor=None):
        """8-bit sRGB (h, w, 3), row 0 = image top; returns (img, e_auto, exposure, anchor)."""
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        img = np.zeros((h, w, 3), np.uint8)
        ea, ex = C.c_double(), C.c_double()
        anc = None if anchor is None else C.c_double(anchor)
        self.lib.or_film_to_rgb8(_ptr(xyz, _dp), w, h, float(n_photons), float(exp_comp), int(bool(absolute)),
                                 C.byref(anc) if anc is not None else None, _ptr(img, _u8p),
                                 C.byref(ea), C.byref(ex))
        return img, ea.value, ex.value, (None if anc is None else anc.value)

    def emission_sample(self, spec, u, pool=None):
        u = np.ascontiguousarray(u, np.float64)
        lam = np.zeros_like(u)
        pdf = np.zeros_like(u)
        integral = C.c_double(0)
        pl = None if pool is None else np.ascontiguousarray(pool, np.float64)
        self._f("emission_sample")(C.byref(spec), None if pl is None else _ptr(pl, _dp), _ptr(u, _dp), len(u),
                                   _ptr(lam, _dp), _ptr(pdf, _dp), C.byref(integral))
        return lam, pdf, integral.value

    def sample_bundle(self, spec, u, c, pool=None):
        lam = np.zeros(8)
        pdf = np.zeros(8)
        pl = None if pool is None else np.ascontiguousarray(pool, np.float64)
        ok = self._f("sample_bundle")(C.byref(spec), None if pl is None else _ptr(pl, _dp), u, c, _ptr(lam, _dp),
                                      _ptr(pdf, _dp))
        return bool(ok), lam[:c], pdf[:c]

    def cosine_hemisphere(self, nrm, seq, seed, n):
        nrm = np.ascontiguousarray(nrm, np.float64)
        out = np.zeros((n, 3))
        self._f("cosine_hemisphere")(_ptr(nrm, _dp), seq, seed, n, _ptr(out, _dp))
        return out

    def sample_glossy(self, mdir, rough, seq, seed, n):
        mdir = np.ascontiguousarray(mdir, np.float64)
        out = np.zeros((n, 3))
        self._f("sample_glossy")(_ptr(mdir, _dp), rough, seq, seed, n, _ptr(out, _dp))
        return out

    def camera_basis(self, cam):
        out = np.zeros(14)
        self._f("camera_basis")(C.byref(cam), _ptr(out, _dp))
        return out

    def camera_project(self, cam, pts):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        n = len(pts)
        px = np.zeros(n, np.int32); py = np.zeros(n, np.int32)
        cc = np.zeros(n); d2 = np.zeros(n); om = np.zeros(n)
        self._f("camera_project")(C.byref(cam), _ptr(pts, _dp), n, _ptr(px, _i32p), _ptr(py, _i32p),
                                  _ptr(cc, _dp), _ptr(d2, _dp), _ptr(om, _dp))
        return px, py, cc, d2, om


class CpuScene:
    def __init__(self, r: _CpuRenderer, sd):
        self.r = r
        self.sd = sd
        self._desc = sd.desc()
        self.h = r._f("scene_create")(C.byref(self._desc))
        if not self.h:
            raise RuntimeError("scene_create failed")

    def close(self):
        if self.h:
            self.r._f("scene_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def bvh_nodes(self):
        return int(self.r._f("scene_bvh_nodes")(self.h))

    @property
    def total_power(self):
        return float(self.r._f("scene_total_power")(self.h))

    def mirror_planes(self):
        """(n, 10) float64: unit normal, offset d, padded box lo and hi of each flat
        mirror plane, in the scene's order."""
        f = self.r._f("scene_mirror_planes")
        n = f(self.h, None, 0)
        out = np.zeros((max(n, 1), 10))
        f(self.h, out.ctypes.data_as(_dp), n)
        return out[:n]

    def closest_hit(self, rays, tmin=1e-6, stats=False):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        n = len(rays)
        prim = np.zeros(n, np.int32)
        t = np.zeros(n)
        nv = np.zeros(n, np.int64)
        lt = np.zeros(n, np.int64)
        self.r._f("closest_hit")(self.h, _ptr(rays, _dp), n, tmin, _ptr(prim, _i32p), _ptr(t, _dp),
                                 _ptr(nv, _i64p), _ptr(lt, _i64p))
        return (prim, t, nv, lt) if stats else (prim, t)

    def closest_hit_linear(self, rays, tmin=1e-6):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        n = len(rays)
        prim = np.zeros(n, np.int32)
        t = np.zeros(n)
        self.r._f("closest_hit_linear")(self.h, _ptr(rays, _dp), n, tmin, _ptr(prim, _i32p), _ptr(t, _dp))
        return prim, t

    def occluded(self, rays, max_dist):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        md = np.ascontiguousarray(max_dist, np.float64)
        occ = np.zeros(len(rays), np.uint8)
        self.r._f("occluded")(self.h, _ptr(rays, _dp), _ptr(md, _dp), len(rays), _ptr(occ, _u8p))
        return occ.astype(bool)


class Oracle(_CpuRenderer):
    prefix = "or_"

    def __init__(self, path: str = ORACLE_SO):
        super().__init__(path)
        L = self.lib
        L.or_render.argtypes = [C.c_void_p, C.POINTER(abi.CameraDesc), C.c_int64, C.c_uint64,
                                C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                C.POINTER(abi.Energy), _u64p]
        L.or_closest_hit_margin.argtypes = [C.c_void_p, _dp, C.c_int64, C.c_double, C.c_double,
                                            _i32p, _dp, _dp, _u8p]
        L.or_closest_hit_ambiguity.argtypes = [C.c_void_p, _dp, _i32p, C.c_int64, C.c_double,
                                               C.c_double, _i32p, _dp, _dp, _u8p, _dp, _dp, _u8p]

    def render(self, scene: CpuScene, cam, n, seed_base=0, salt=abi.REFERENCE_SALT, hero_c=4,
               max_bounce=32, rr_depth=0, threads=1):
        npx = cam.res_x * cam.res_y
        xyz = np.zeros((npx, 3))
        hits = np.zeros(npx)
        e = abi.Energy()
        st = np.zeros(7, np.uint64)
        self.lib.or_render(scene.h, C.byref(cam), n, seed_base, salt, hero_c, max_bounce, rr_depth,
                           threads, _ptr(xyz, _dp), _ptr(hits, _dp), C.byref(e), _ptr(st, _u64p))
        keys = ("segments", "connections", "splats", "node_visits", "leaf_tests", "any_node_visits",
                "any_leaf_tests")
        return xyz, hits, e.as_dict(), {k: int(v) for k, v in zip(keys, st)}

    def closest_hit_margin(self, scene: CpuScene, rays, tmin=1e-6, rel_margin=1e-5):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        n = len(rays)
        prim = np.zeros(n, np.int32)
        t = np.zeros(n)
        t2 = np.zeros(n)
        edge = np.zeros(n, np.uint8)
        self.lib.or_closest_hit_margin(scene.h, _ptr(rays, _dp), n, tmin, rel_margin,
                                       _ptr(prim, _i32p), _ptr(t, _dp), _ptr(t2, _dp), _ptr(edge, _u8p))
        return prim, t, t2, edge.astype(bool)

    def closest_hit_ambiguity(self, scene: CpuScene, rays, query, tmin=1e-6, rel_margin=1e-5):
        """Brute-force ambiguity report per ray (or_closest_hit_ambiguity): the exact
        closest hit, the nearest grown-hit distance of another triangle, whether the
        winner is grazed, the nearest robust hit, and the grown-hit distance and graze
        flag of the queried triangle query[i] (-1 = none)."""
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        q = np.ascontiguousarray(query, np.int32)
        n = len(rays)
        assert len(q) == n
        prim = np.zeros(n, np.int32)
        t, t2, trob, tq = (np.zeros(n) for _ in range(4))
        edge = np.zeros(n, np.uint8)
        qg = np.zeros(n, np.uint8)
        self.lib.or_closest_hit_ambiguity(scene.h, _ptr(rays, _dp), _ptr(q, _i32p), n, tmin, rel_margin,
                                          _ptr(prim, _i32p), _ptr(t, _dp), _ptr(t2, _dp), _ptr(edge, _u8p),
                                          _ptr(trob, _dp), _ptr(tq, _dp), _ptr(qg, _u8p))
        return {"prim": prim, "t": t, "t2": t2, "edge": edge.astype(bool), "t_robust": trob, "tq": tq,
                "q_graze": qg.astype(bool)}


class Reference(_CpuRenderer):
    prefix = "ref_"

    def __init__(self, path: str = REF_SO):
        super().__init__(path)
        self.lib.ref_render.argtypes = [C.c_void_p, C.POINTER(abi.CameraDesc), C.c_int64,
                                        C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, _dp,
                                        _dp, C.POINTER(abi.Energy), _u64p]

    def render(self, scene: CpuScene, cam, n, seed_base=0, salt=abi.REFERENCE_SALT, hero_c=4,
               max_bounce=32, threads=1):
        npx = cam.res_x * cam.res_y
        xyz = np.zeros((npx, 3))
        hits = np.zeros(npx)
        e = abi.Energy()
        rq = C.c_uint64(0)
        self.lib.ref_render(scene.h, C.byref(cam), n, seed_base, salt, hero_c, max_bounce, threads,
                            _ptr(xyz, _dp), _ptr(hits, _dp), C.byref(e), C.byref(rq))
        return xyz, hits, e.as_dict(), {"ray_queries": int(rq.value)}


def available(which: str = "oracle") -> bool:
    return os.path.exists({"oracle": ORACLE_SO, "ref": REF_SO, "adapter": ADAPTER_SO}[which])


ADAPTER_SO = os.path.join(HERE, "_ref", "libref_adapter.so")


class AdapterScene:
    """A scene built with the reference's own API inside oracle/_ref/libref_adapter.so
    (ref_adapter_check.cpp), seen through the drop-in adapter
    (goodphotons_amd/host/ref_adapter.hpp). `which`: 0 box, 1 box + glass + glossy +
    mirror, 10..16 variants outside the supported set."""

    def __init__(self, lib, which):
        self.L = lib
        self.which = which
        self.h = lib.adapter_scene(which)

    def supported(self):
        buf = C.create_string_buffer(256)
        ok = self.L.adapter_supported(self.h, buf, 256)
        return bool(ok), buf.value.decode()

    def desc(self):
        """The gp_scene_desc the adapter hands over (a view into the adapter's memory)."""
        p = self.L.adapter_desc(self.h)
        if not p:
            raise RuntimeError("scene outside the supported set")
        return abi.SceneDesc.from_address(p)

    @property
    def n_tris(self):
        return int(self.L.adapter_n_tris(self.h))

    def close(self):
        if self.h:
            self.L.adapter_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefAdapter:
    """ctypes wrapper of oracle/_ref/libref_adapter.so."""

    def __init__(self, path: str = ADAPTER_SO):
        L = self.lib = C.CDLL(path)
        L.adapter_scene.restype = C.c_void_p
        L.adapter_scene.argtypes = [C.c_int]
        L.adapter_scene_free.argtypes = [C.c_void_p]
        L.adapter_supported.restype = C.c_int
        L.adapter_supported.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.adapter_desc.restype = C.c_void_p
        L.adapter_desc.argtypes = [C.c_void_p]
        L.adapter_n_tris.restype = C.c_int64
        L.adapter_n_tris.argtypes = [C.c_void_p]
        L.adapter_camera_desc.argtypes = [C.c_int, C.POINTER(abi.CameraDesc)]
        L.adapter_render_gpu.argtypes = [C.c_void_p, C.c_int, C.c_longlong, C.c_ulonglong, C.c_int, C.c_int,
                                         C.c_int, _dp, _dp, _dp]
        L.adapter_render_reference.argtypes = [C.c_void_p, C.c_int, C.c_longlong, C.c_ulonglong, C.c_int,
                                               C.c_int, C.c_int, _dp, _dp, _dp]
        L.adapter_shared_session.restype = C.c_int
        L.adapter_shared_session.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_longlong, C.c_int, C.c_int, C.c_int,
                                             C.c_int] + [_dp] * 7

    def scene(self, which) -> AdapterScene:
        return AdapterScene(self.lib, which)

    def camera(self, res):
        c = abi.CameraDesc()
        self.lib.adapter_camera_desc(res, C.byref(c))
        return c

    def _render(self, fn, s, res, n, seed_base, hero_c, depth, extra):
        xyz = np.zeros((res * res, 3))
        hits = np.zeros(res * res)
        e = np.zeros(5)
        fn(s.h, res, n, seed_base, hero_c, depth, extra, _ptr(xyz, _dp), _ptr(hits, _dp), _ptr(e, _dp))
        keys = ("emitted", "absorbed", "sensor", "escaped", "residual")
        return xyz, hits, dict(zip(keys, e))

    def render_gpu(self, s, res, n, seed_base=0, hero_c=4, depth=8, bvh=2):
        """gpref::renderForward (the HIP library); bvh: GP_BVH_PREBUILT (2, the scene's
        own tree) or GP_BVH_HOST_SAH (0)."""
        return self._render(self.lib.adapter_render_gpu, s, res, n, seed_base, hero_c, depth, bvh)

    def render_reference(self, s, res, n, seed_base=0, hero_c=4, depth=8, threads=1):
        return self._render(self.lib.adapter_render_reference, s, res, n, seed_base, hero_c, depth, threads)

    def shared_session(self, s, res, res2, n, batches, k, hero_c=4, depth=8):
        """gpref::SharedSession over two cameras: (films, hits0, energy) uninterrupted and
        (films, energy) stopped after k batches and resumed from the downloaded totals;
        films = [(xyz, hits)] per camera."""
        a = [np.zeros(res * res * 4), np.zeros(res2 * res2 * 4)]
        b = [np.zeros(res * res * 4), np.zeros(res2 * res2 * 4)]
        h0 = np.zeros(res * res)
        e, re = np.zeros(5), np.zeros(5)
        ok = self.lib.adapter_shared_session(s.h, res, res2, n, batches, k, hero_c, depth, _ptr(a[0], _dp),
                                             _ptr(a[1], _dp), _ptr(h0, _dp), _ptr(e, _dp), _ptr(b[0], _dp),
                                             _ptr(b[1], _dp), _ptr(re, _dp))
        if not ok:
            raise RuntimeError("gpref::SharedSession::begin failed")
        split = lambda v, r: (v[:3 * r * r].reshape(-1, 3), v[3 * r * r:])
        return ([split(a[0], res), split(a[1], res2)], h0, e, [split(b[0], res), split(b[1], res2)], re)
