"""MI355X-native forward light-tracing core (host side, over the C ABI).

Mirrors the reference renderer's GPU forward entry point
(`renderForwardCuda`, src/render_cuda.h:79-82): a scene and camera go in, an
XYZ film (un-normalised sums, per-pixel hit counts) and an accumulated energy
ledger come back. The work runs in libgoodphotons.so (hand-written HIP kernels
for gfx950); there is no CPU fallback: if the library or a gfx950 device is
missing, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

__all__ = ["Library", "GpuScene", "Film", "Server", "render_forward", "GpError", "load"]


class GpError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_LIB = None


class Library:
    """The loaded HIP library with typed entry points."""

    def __init__(self, path: str = abi.LIB_PATH):
        if not os.path.exists(path):
            raise GpError(3, f"{path} is not built; run goodphotons_amd.build.build()")
        self.path = path
        L = self.lib = C.CDLL(path)
        vp = C.c_void_p
        L.gp_abi_version.restype = C.c_int32
        L.gp_status_string.restype = C.c_char_p
        L.gp_status_string.argtypes = [C.c_int]
        L.gp_last_error.restype = C.c_char_p
        L.gp_device_query.argtypes = [C.POINTER(abi.DeviceInfo)]
        L.gp_scene_create.argtypes = [C.POINTER(abi.SceneDesc), C.POINTER(vp)]
        L.gp_scene_create_ex.argtypes = [C.POINTER(abi.SceneDesc), C.c_int32, C.POINTER(vp)]
        L.gp_scene_get_info.argtypes = [vp, C.POINTER(abi.SceneInfo)]
        L.gp_scene_destroy.argtypes = [vp]
        L.gp_scene_destroy.restype = None
        L.gp_mirror_planes.restype = C.c_int32
        L.gp_mirror_planes.argtypes = [C.POINTER(abi.SceneDesc), C.POINTER(C.c_double), C.c_int32]
        L.gp_film_create.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.POINTER(vp)]
        L.gp_film_reset.argtypes = [vp]
        L.gp_film_download.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gp_film_download_raw.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.gp_film_download_hits.argtypes = [vp, C.POINTER(C.c_double)]
        L.gp_film_seed.argtypes = [vp, vp, C.POINTER(abi.CameraDesc), C.POINTER(abi.TraceParams),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gp_film_destroy.argtypes = [vp]
        L.gp_film_destroy.restype = None
        L.gp_film_to_linear.argtypes = [vp, C.c_int64, C.POINTER(C.c_float)]
        L.gp_checkpoint_guard.restype = C.c_uint64
        L.gp_checkpoint_guard.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_int32,
                                          C.c_int32, C.c_int32]
        L.gp_film_save_checkpoint.argtypes = [vp, C.c_char_p, C.c_int64, C.POINTER(abi.Energy), C.c_int32,
                                              C.c_uint64]
        L.gp_film_load_checkpoint.argtypes = [C.c_char_p, C.c_int32, C.c_uint64, C.POINTER(vp),
                                              C.POINTER(C.c_int64), C.POINTER(abi.Energy)]
        L.gp_film_tonemap.argtypes = [vp, C.c_int64, C.c_double, C.c_int32, C.POINTER(C.c_double),
                                      C.POINTER(C.c_uint8), C.POINTER(C.c_double)]
        L.gp_film_tonemap_noalloc.argtypes = L.gp_film_tonemap.argtypes
        L.gp_trace.argtypes = [vp, C.POINTER(abi.CameraDesc), vp, C.POINTER(abi.TraceParams),
                               C.POINTER(abi.Energy), C.POINTER(abi.TraceStats)]
        L.gp_trace_shared.argtypes = [vp, C.c_int32, C.POINTER(abi.CameraDesc), C.POINTER(vp),
                                      C.POINTER(abi.TraceParams), C.POINTER(abi.Energy), C.POINTER(abi.TraceStats)]
        L.gp_probe_closest_hit.argtypes = [vp, C.POINTER(C.c_float), C.c_int64, C.c_float,
                                           C.POINTER(C.c_int32), C.POINTER(C.c_float)]
        L.gp_probe_any_hit.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64,
                                       C.c_float, C.POINTER(C.c_uint8)]
        L.gp_get_tuning.argtypes = [C.POINTER(abi.Tuning)]
        L.gp_set_tuning.argtypes = [C.POINTER(abi.Tuning)]
        L.gp_autotune.argtypes = [vp, C.POINTER(abi.CameraDesc), C.POINTER(abi.TraceParams),
                                  C.POINTER(abi.Tuning)]
        L.gp_server_create.argtypes = [C.POINTER(abi.ServerConfig), C.POINTER(vp)]
        L.gp_server_submit.argtypes = [vp, C.POINTER(abi.JobDesc), C.POINTER(C.c_int64)]
        L.gp_server_step.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
        L.gp_server_take_result.argtypes = [vp, C.c_int64, C.POINTER(abi.JobResult),
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gp_server_poll.argtypes = [vp, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32)]
        L.gp_server_pending.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.gp_server_destroy.argtypes = [vp]
        L.gp_server_destroy.restype = None

    def check(self, status: int):
        if status != 0:
            msg = self.lib.gp_last_error()
            raise GpError(status, msg.decode() if msg else "")

    def mirror_planes(self, sd) -> np.ndarray:
        """(n, 10) float64: the flat mirror planes gp_scene_create derives from the
        scene (unit normal, offset d, padded box lo and hi). Host only."""
        d = sd.desc()
        n = self.lib.gp_mirror_planes(C.byref(d), None, 0)
        if n < 0:
            self.check(-n)
        out = np.zeros((max(n, 1), 10))
        self.lib.gp_mirror_planes(C.byref(d), out.ctypes.data_as(C.POINTER(C.c_double)), n)
        return out[:n]

    def device_info(self) -> dict:
        d = abi.DeviceInfo()
        self.check(self.lib.gp_device_query(C.byref(d)))
        return {"name": d.name.decode(), "arch": d.arch.decode(), "compute_units": d.compute_units,
                "total_mem": d.total_mem, "free_mem": d.free_mem, "lds_per_block": d.lds_per_block}

    def tuning(self) -> abi.Tuning:
        t = abi.Tuning()
        self.check(self.lib.gp_get_tuning(C.byref(t)))
        return t

    def set_tuning(self, **kw) -> abi.Tuning:
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        self.check(self.lib.gp_set_tuning(C.byref(t)))
        return t


def load() -> Library:
    global _LIB
    if _LIB is None:
        _LIB = Library()
    return _LIB


class GpuScene:
    """A scene resident in HBM (BVH, triangles, materials, emitter tables)."""

    BUILDERS = {"host": 0, "device": 1}

    def __init__(self, sd, lib: Library | None = None, builder: str = "host"):
        """builder: "host" (binned SAH, gp_scene_create) or "device" (linear BVH built
        on the GPU, gp_scene_create_ex)."""
        self.L = lib or load()
        self.sd = sd
        self._desc = sd.desc()
        h = C.c_void_p()
        self.L.check(self.L.lib.gp_scene_create_ex(C.byref(self._desc), self.BUILDERS[builder], C.byref(h)))
        self.h = h

    def info(self) -> dict:
        i = abi.SceneInfo()
        self.L.check(self.L.lib.gp_scene_get_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    def closest_hit(self, rays, tmin=1e-6):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = len(rays)
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        self.L.check(self.L.lib.gp_probe_closest_hit(
            self.h, rays.ctypes.data_as(C.POINTER(C.c_float)), n, tmin,
            prim.ctypes.data_as(C.POINTER(C.c_int32)), t.ctypes.data_as(C.POINTER(C.c_float))))
        return prim, t

    def any_hit(self, rays, tmax, tmin=1e-6):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        tm = np.ascontiguousarray(tmax, np.float32)
        occ = np.zeros(len(rays), np.uint8)
        self.L.check(self.L.lib.gp_probe_any_hit(
            self.h, rays.ctypes.data_as(C.POINTER(C.c_float)), tm.ctypes.data_as(C.POINTER(C.c_float)),
            len(rays), tmin, occ.ctypes.data_as(C.POINTER(C.c_uint8))))
        return occ.astype(bool)

    def close(self):
        if getattr(self, "h", None):
            self.L.lib.gp_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Film:
    """int64 fixed-point XYZ + hits film in HBM."""

    def __init__(self, res_x, res_y, max_photons=0, lib: Library | None = None):
        self.L = lib or load()
        self.rx, self.ry = int(res_x), int(res_y)
        h = C.c_void_p()
        self.L.check(self.L.lib.gp_film_create(self.rx, self.ry, int(max_photons), C.byref(h)))
        self.h = h

    def download(self):
        npx = self.rx * self.ry
        xyz = np.zeros((npx, 3))
        hits = np.zeros(npx)
        self.L.check(self.L.lib.gp_film_download(self.h, xyz.ctypes.data_as(C.POINTER(C.c_double)),
                                                 hits.ctypes.data_as(C.POINTER(C.c_double))))
        return xyz, hits

    def raw(self):
        w = np.zeros((self.rx * self.ry, 4), np.int64)
        sc = C.c_double(0)
        self.L.check(self.L.lib.gp_film_download_raw(self.h, w.ctypes.data_as(C.POINTER(C.c_int64)),
                                                     C.byref(sc)))
        return w, sc.value

    def download_hits(self):
        """Hit counts only (gp_film_download_hits)."""
        hits = np.zeros(self.rx * self.ry)
        self.L.check(self.L.lib.gp_film_download_hits(self.h, hits.ctypes.data_as(C.POINTER(C.c_double))))
        return hits

    def seed(self, scene: "GpuScene", cam, n_photons, xyz, hits, max_depth=32, hero_c=4):
        """gp_film_seed: running totals (doubles) of n_photons photons into this fresh film."""
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        h = np.ascontiguousarray(hits, np.float64).reshape(-1)
        p = trace_params(n_photons, hero_c=hero_c, max_depth=max_depth)
        self.L.check(self.L.lib.gp_film_seed(self.h, scene.h, C.byref(cam), C.byref(p),
                                             x.ctypes.data_as(C.POINTER(C.c_double)),
                                             h.ctypes.data_as(C.POINTER(C.c_double))))

    def reset(self):
        self.L.check(self.L.lib.gp_film_reset(self.h))

    def save_checkpoint(self, path, n_photons, energy=None, mode="B", guard=0):
        """FTBUF01 checkpoint (+ exact fixed-point section) of this film."""
        e = energy if isinstance(energy, abi.Energy) else None
        self.L.check(self.L.lib.gp_film_save_checkpoint(self.h, os.fsencode(path), int(n_photons),
                                                        C.byref(e) if e is not None else None,
                                                        ord(mode), int(guard)))

    @classmethod
    def load_checkpoint(cls, path, mode="B", guard=0, lib: "Library | None" = None):
        """(film, cumulative photons, energy) from a checkpoint written by save_checkpoint."""
        L = lib or load()
        h = C.c_void_p()
        n = C.c_int64(0)
        e = abi.Energy()
        L.check(L.lib.gp_film_load_checkpoint(os.fsencode(path), ord(mode), int(guard), C.byref(h),
                                              C.byref(n), C.byref(e)))
        with open(path, "rb") as fh:
            rx, ry = np.frombuffer(fh.read(16)[8:16], "<i4")
        f = cls.__new__(cls)
        f.L, f.h, f.rx, f.ry = L, h, int(rx), int(ry)
        return f, n.value, e

    def to_linear(self, n_photons):
        """Scene-linear sRGB, float32 (res_y, res_x, 3), row 0 = bottom (PFM order)."""
        out = np.zeros((self.ry, self.rx, 3), np.float32)
        self.L.check(self.L.lib.gp_film_to_linear(self.h, int(n_photons),
                                                  out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def tonemap(self, n_photons, exposure_comp=0.0, absolute=False, anchor=None, no_alloc=False):
        """8-bit sRGB (res_y, res_x, 3), row 0 = top, with the p99 auto exposure.
        Returns (image, exposure, anchor); pass the returned anchor back to lock it.
        no_alloc=True uses gp_film_tonemap_noalloc (no device allocation; films up to
        2048 x 2048)."""
        img = np.zeros((self.ry, self.rx, 3), np.uint8)
        ex = C.c_double(0)
        anc = C.c_double(0.0 if anchor is None else float(anchor))
        fn = self.L.lib.gp_film_tonemap_noalloc if no_alloc else self.L.lib.gp_film_tonemap
        self.L.check(fn(self.h, int(n_photons), float(exposure_comp), int(bool(absolute)), C.byref(anc),
                        img.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(ex)))
        return img, ex.value, anc.value

    def close(self):
        if getattr(self, "h", None):
            self.L.lib.gp_film_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_pfm(path, lin):
    """PFM (writePfm, main.cpp:10372-10387): 'PF', width height, -1.0 (little endian),
    then float32 RGB rows bottom to top -- the row order of `to_linear`."""
    lin = np.ascontiguousarray(lin, np.float32)
    h, w = lin.shape[:2]
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(lin.astype("<f4").tobytes())


def write_ppm(path, rgb8):
    """Binary PPM of an 8-bit image whose row 0 is the top."""
    rgb8 = np.ascontiguousarray(rgb8, np.uint8)
    h, w = rgb8.shape[:2]
    with open(path, "wb") as f:
        f.write(f"P6\n{w} {h}\n255\n".encode())
        f.write(rgb8.tobytes())


def write_png(path, rgb8):
    """8-bit RGB PNG (zlib-compressed scanlines, filter 0) of an image whose row 0 is the top."""
    import struct
    import zlib
    rgb8 = np.ascontiguousarray(rgb8, np.uint8)
    h, w = rgb8.shape[:2]
    raw = b"".join(b"\x00" + rgb8[y].tobytes() for y in range(h))

    def chunk(t, d):
        c = struct.pack(">I", len(d)) + t + d
        return c + struct.pack(">I", zlib.crc32(t + d) & 0xffffffff)
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))


def trace_params(n, seed_base=0, salt=abi.REFERENCE_SALT, hero_c=4, max_depth=32, rr_depth=0,
                 traversal_stats=False):
    p = abi.TraceParams()
    p.n_photons = int(n)
    p.seed_base = int(seed_base)
    p.salt = int(salt)
    p.hero_c = int(hero_c)
    p.max_depth = int(max_depth)
    p.rr_depth = int(rr_depth)
    p.flags = abi.TRACE_TRAVERSAL_STATS if traversal_stats else 0
    return p


def trace(scene: GpuScene, cam, film: Film, n, seed_base=0, salt=abi.REFERENCE_SALT, hero_c=4,
          max_depth=32, rr_depth=0, energy: abi.Energy | None = None, traversal_stats=False):
    """gp_trace; traversal_stats=True also counts node visits and triangle tests
    (GP_TRACE_TRAVERSAL_STATS) into the returned stats."""
    p = trace_params(n, seed_base, salt, hero_c, max_depth, rr_depth, traversal_stats)
    e = energy if energy is not None else abi.Energy()
    st = abi.TraceStats()
    scene.L.check(scene.L.lib.gp_trace(scene.h, C.byref(cam), film.h, C.byref(p), C.byref(e), C.byref(st)))
    return e, st


def render_forward(scene: GpuScene, cam, n, seed_base=0, hero_c=4, max_depth=32, rr_depth=0,
                   salt=abi.REFERENCE_SALT):
    """Trace n photons into a fresh film: returns (xyz[res,3], hits[res], energy, stats).

    Same contract as renderForwardCuda: a new film each call, seed_base = photons
    already traced by earlier calls, energy returned for accumulation by the caller.
    """
    film = Film(cam.res_x, cam.res_y, max_photons=max(1, int(n)), lib=scene.L)
    try:
        e, st = trace(scene, cam, film, n, seed_base, salt, hero_c, max_depth, rr_depth)
        xyz, hits = film.download()
    finally:
        film.close()
    return xyz, hits, e.as_dict(), st.as_dict()


def trace_shared(scene: GpuScene, cams, films, n, seed_base=0, salt=abi.REFERENCE_SALT, hero_c=4,
                 max_depth=32, rr_depth=0, energy: abi.Energy | None = None):
    """One photon flight splatted into every camera (gp_trace_shared)."""
    p = trace_params(n, seed_base, salt, hero_c, max_depth, rr_depth)
    e = energy if energy is not None else abi.Energy()
    st = abi.TraceStats()
    ca = (abi.CameraDesc * len(cams))(*cams)
    fa = (C.c_void_p * len(films))(*[f.h.value for f in films])
    scene.L.check(scene.L.lib.gp_trace_shared(scene.h, len(cams), ca, fa, C.byref(p), C.byref(e), C.byref(st)))
    return e, st


def render_forward_shared(scene: GpuScene, cams, n, seed_base=0, hero_c=4, max_depth=32, rr_depth=0,
                          salt=abi.REFERENCE_SALT):
    """renderForwardSharedCuda: fresh films, one per camera; returns ([(xyz, hits)], energy, stats)."""
    films = [Film(c.res_x, c.res_y, max_photons=max(1, int(n)), lib=scene.L) for c in cams]
    try:
        e, st = trace_shared(scene, cams, films, n, seed_base, salt, hero_c, max_depth, rr_depth)
        out = [f.download() for f in films]
    finally:
        for f in films:
            f.close()
    return out, e.as_dict(), st.as_dict()


class Server:
    """Continuous-batching render server over resident scenes (gp_server_*)."""

    def __init__(self, hbm_budget_bytes=16 << 30, max_live_jobs=64, lib: Library | None = None,
                 order_policy=abi.ORDER_SHORTEST_FIRST):
        self.L = lib or load()
        cfg = abi.ServerConfig()
        cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        cfg.max_live_jobs = int(max_live_jobs)
        cfg.order_policy = int(order_policy)
        h = C.c_void_p()
        self.L.check(self.L.lib.gp_server_create(C.byref(cfg), C.byref(h)))
        self.h = h

    def submit(self, scene: GpuScene, cam, n, seed_base=0, salt=abi.REFERENCE_SALT, hero_c=4,
               max_depth=16, rr_depth=0, tag=0) -> int:
        j = abi.JobDesc()
        j.scene = scene.h.value
        j.camera = cam
        j.params = trace_params(n, seed_base, salt, hero_c, max_depth, rr_depth)
        j.user_tag = int(tag)
        jid = C.c_int64(0)
        self.L.check(self.L.lib.gp_server_submit(self.h, C.byref(j), C.byref(jid)))
        return jid.value

    def step(self, max_iterations=0) -> int:
        n = C.c_int32(0)
        self.L.check(self.L.lib.gp_server_step(self.h, int(max_iterations), C.byref(n)))
        return n.value

    def poll(self, capacity=256):
        """Ids of finished jobs not yet taken, oldest first."""
        ids = (C.c_int64 * capacity)()
        n = C.c_int32(0)
        self.L.check(self.L.lib.gp_server_poll(self.h, ids, int(capacity), C.byref(n)))
        return [ids[i] for i in range(n.value)]

    def pending(self, with_bytes=False):
        """(queued, live) jobs, and with_bytes the footprint the live ones hold."""
        q, l, b = C.c_int32(0), C.c_int32(0), C.c_int64(0)
        self.L.check(self.L.lib.gp_server_pending(self.h, C.byref(q), C.byref(l), C.byref(b)))
        return (q.value, l.value, b.value) if with_bytes else (q.value, l.value)

    def take(self, job_id, res_x, res_y, film=True):
        r = abi.JobResult()
        xyz = np.zeros((res_x * res_y, 3)) if film else None
        hits = np.zeros(res_x * res_y) if film else None
        self.L.check(self.L.lib.gp_server_take_result(
            self.h, int(job_id), C.byref(r),
            xyz.ctypes.data_as(C.POINTER(C.c_double)) if film else None,
            hits.ctypes.data_as(C.POINTER(C.c_double)) if film else None))
        return r, xyz, hits

    def close(self):
        if getattr(self, "h", None):
            self.L.lib.gp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
