"""Seeded synthetic scenes for the forward light-tracing workloads.

Every scene is a triangle list (float32 vertices, so the GPU library, the oracle
and the reference harness all consume bit-identical geometry), a material table
and quad emitters, plus a pinhole camera. Builders:

  cornell_c0        Cornell-style box: 5 wall quads (10 tris), two diffuse boxes
                    (12 tris each), a 0.25 x 0.25 ceiling light quad (2 tris),
                    6500 K blackbody band-limited to 380-780 nm, 256^2 film.
  caustics_c1       the c0 box and boxes + 64x64-segment UV glass sphere (N-BK7
                    Sellmeier) + triangular prism (SF10 Sellmeier) + a small
                    0.04 x 0.04 quad light ("0.02 radius") aimed down through
                    both, 512^2 film.
  icosphere_c2      icosphere subdivided `level` times (20*4^level tris; 9 gives
                    5,242,880), radially displaced by seeded 3-octave value noise of
                    amplitude 0.05, dielectric n = 1.5 + 0.004/um^2, on a diffuse
                    floor, 3200 K quad light, 1024^2 film.
  frames_c3         animation frame f of c1: sphere rotated 2*pi*f/64 about y,
                    prism translated along x by 0.1*sin(2*pi*f/64).
  soup              seeded random-triangle soup in the unit cube, edge lengths
                    uniform in [0.002, 0.02], diffuse albedo uniform in [0.2, 0.9]
                    (quantised to 1024 levels, one material per level).

Geometry conventions follow the reference's Cornell builder (src/main.cpp:329-381):
quads are split (a,b,c),(a,c,d); the camera matches scenes/cornell.ftsl:38-42.
Closed dielectric meshes are wound outward, because the reference decides
entering/exiting from the triangle's geometric normal (src/render.h:2661).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

# Schott catalogue Sellmeier coefficients (lambda in um, C in um^2).
BK7_SELLMEIER = (1.03961212, 0.231792344, 1.01046945, 0.00600069867, 0.0200179144, 103.560653)
SF10_SELLMEIER = (1.62153902, 0.256287842, 1.64447552, 0.0122241457, 0.0595736775, 147.468793)


@dataclass
class SceneData:
    name: str
    verts: np.ndarray                      # (n, 9) float32
    mat_ids: np.ndarray                    # (n,) int32
    materials: list = field(default_factory=list)   # list[abi.Material]
    emitters: list = field(default_factory=list)    # list[abi.Emitter]
    camera: abi.CameraDesc | None = None
    n_photons: int = 1 << 20
    max_depth: int = 16
    rr_depth: int = 0
    hero_c: int = 4
    spectrum_samples: np.ndarray | None = None   # samples of SPEC_SAMPLED spectra

    @property
    def n_tris(self) -> int:
        return int(self.verts.shape[0])

    def desc(self) -> abi.SceneDesc:
        """ctypes descriptor; the arrays it points to are kept alive on self."""
        self._v = np.ascontiguousarray(self.verts, dtype=np.float32)
        self._m = np.ascontiguousarray(self.mat_ids, dtype=np.int32)
        self._mats = (abi.Material * max(1, len(self.materials)))(*self.materials)
        self._ems = (abi.Emitter * max(1, len(self.emitters)))(*self.emitters)
        d = abi.SceneDesc()
        d.n_tris = self.n_tris
        d.n_materials = len(self.materials)
        d.n_emitters = len(self.emitters)
        d.verts = self._v.ctypes.data_as(C.POINTER(C.c_float))
        d.mat_ids = self._m.ctypes.data_as(C.POINTER(C.c_int32))
        d.materials = C.cast(self._mats, C.POINTER(abi.Material))
        d.emitters = C.cast(self._ems, C.POINTER(abi.Emitter))
        if self.spectrum_samples is not None and len(self.spectrum_samples):
            self._samples = np.ascontiguousarray(self.spectrum_samples, dtype=np.float64)
            d.spectrum_samples = self._samples.ctypes.data_as(C.POINTER(C.c_double))
            d.n_spectrum_samples = len(self._samples)
        self._desc = d
        return d

    def with_camera_res(self, rx: int, ry: int | None = None) -> "SceneData":
        cam = camera(tuple(self.camera.eye), tuple(self.camera.target), tuple(self.camera.up),
                     self.camera.fov_y_deg, rx, ry if ry is not None else rx)
        out = SceneData(self.name, self.verts, self.mat_ids, list(self.materials),
                        list(self.emitters), cam, self.n_photons, self.max_depth,
                        self.rr_depth, self.hero_c, self.spectrum_samples)
        return out


# ---------------------------------------------------------------------------- helpers
def diffuse(spec: abi.Spectrum) -> abi.Material:
    m = abi.Material()
    m.type = abi.MAT_DIFFUSE
    m.roughness = 0.1
    m.reflect = spec
    m.ior = abi.spectrum(abi.SPEC_CONST, 1.5)
    return m


def dielectric(ior: abi.Spectrum, roughness: float = 0.0) -> abi.Material:
    m = abi.Material()
    m.type = abi.MAT_DIELECTRIC
    m.roughness = roughness
    m.reflect = abi.spectrum(abi.SPEC_CONST, 0.5)
    m.ior = ior
    return m


def glossy(spec: abi.Spectrum, roughness: float) -> abi.Material:
    m = abi.Material()
    m.type = abi.MAT_GLOSSY
    m.roughness = roughness
    m.reflect = spec
    m.ior = abi.spectrum(abi.SPEC_CONST, 1.5)
    return m


def mirror(spec: abi.Spectrum) -> abi.Material:
    """Perfect mirror (delta reflection) with spectral reflectance `spec`."""
    m = abi.Material()
    m.type = abi.MAT_MIRROR
    m.roughness = 0.0
    m.reflect = spec
    m.ior = abi.spectrum(abi.SPEC_CONST, 1.5)
    return m


def quad_emitter(origin, u, v, normal, spd: abi.Spectrum, collimated=False, beam=(1, 0, 0)):
    e = abi.Emitter()
    abi.vec3(e.origin, origin)
    abi.vec3(e.u, u)
    abi.vec3(e.v, v)
    abi.vec3(e.normal, normal)
    e.area = float(np.linalg.norm(np.cross(np.asarray(u, float), np.asarray(v, float))))
    e.collimated = 1 if collimated else 0
    abi.vec3(e.beam_dir, beam)
    e.spd = spd
    return e


def camera(eye, target, up, fov_y, rx, ry) -> abi.CameraDesc:
    c = abi.CameraDesc()
    abi.vec3(c.eye, eye)
    abi.vec3(c.target, target)
    abi.vec3(c.up, up)
    c.fov_y_deg = float(fov_y)
    c.res_x, c.res_y = int(rx), int(ry)
    return c


class _Builder:
    def __init__(self):
        self.tris: list[np.ndarray] = []
        self.mats: list[np.ndarray] = []

    def quad(self, a, b, c, d, mat):
        self.tri(a, b, c, mat)
        self.tri(a, c, d, mat)

    def tri(self, a, b, c, mat):
        self.tris.append(np.array([a, b, c], dtype=np.float64).reshape(1, 9))
        self.mats.append(np.array([mat], dtype=np.int32))

    def mesh(self, tris9: np.ndarray, mat):
        tris9 = np.asarray(tris9, dtype=np.float64).reshape(-1, 9)
        self.tris.append(tris9)
        if np.isscalar(mat):
            self.mats.append(np.full(tris9.shape[0], mat, dtype=np.int32))
        else:
            self.mats.append(np.asarray(mat, dtype=np.int32))

    def arrays(self):
        v = np.concatenate(self.tris, axis=0).astype(np.float32)
        m = np.concatenate(self.mats, axis=0).astype(np.int32)
        return v, m


def _box_tris(center_xz, size, height, angle_deg):
    """Closed box (12 tris, outward winding) standing on y=0, rotated about y."""
    cx, cz = center_xz
    sx, sz = size
    th = math.radians(angle_deg)
    ca, sa = math.cos(th), math.sin(th)

    def P(x, y, z):
        x -= 0.0
        return (cx + ca * x + sa * z, y, cz - sa * x + ca * z)

    hx, hz = sx / 2, sz / 2
    c = [P(-hx, 0, -hz), P(hx, 0, -hz), P(hx, 0, hz), P(-hx, 0, hz),
         P(-hx, height, -hz), P(hx, height, -hz), P(hx, height, hz), P(-hx, height, hz)]
    quads = [(0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0)]
    out = []
    for a, b, cc, d in quads:
        out.append([c[a], c[b], c[cc]])
        out.append([c[a], c[cc], c[d]])
    arr = np.array(out, dtype=np.float64).reshape(-1, 9)
    return _orient_outward(arr, np.array([cx, height / 2, cz]))


def _orient_outward(tris9: np.ndarray, center: np.ndarray) -> np.ndarray:
    """Flip triangles whose geometric normal points toward `center` (convex solids)."""
    t = tris9.reshape(-1, 3, 3).copy()
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    c = t.mean(axis=1) - center
    flip = np.einsum("ij,ij->i", n, c) < 0
    t[flip, 1], t[flip, 2] = t[flip, 2].copy(), t[flip, 1].copy()
    return t.reshape(-1, 9)


def _uv_sphere(center, radius, nseg=64, nring=64, rot_y=0.0):
    """UV sphere with `nseg` longitudes and `nring` latitude bands, fan caps at the
    poles (2*nseg*(nring-1) tris; the 128 zero-area pole-quad halves of a naive
    64x64x2 split are not emitted). Outward winding."""
    verts = [(0.0, 1.0, 0.0)]
    for r in range(1, nring):
        th = math.pi * r / nring
        for s in range(nseg):
            ph = 2 * math.pi * s / nseg + rot_y
            verts.append((math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)))
    verts.append((0.0, -1.0, 0.0))
    V = np.array(verts) * radius + np.asarray(center, float)
    south = len(verts) - 1

    def ring(r, s):
        return 1 + (r - 1) * nseg + (s % nseg)

    faces = []
    for s in range(nseg):
        faces.append((0, ring(1, s + 1), ring(1, s)))
    for r in range(1, nring - 1):
        for s in range(nseg):
            a, b = ring(r, s), ring(r, s + 1)
            c, d = ring(r + 1, s + 1), ring(r + 1, s)
            faces.append((a, b, c))
            faces.append((a, c, d))
    for s in range(nseg):
        faces.append((south, ring(nring - 1, s), ring(nring - 1, s + 1)))
    F = np.array(faces)
    tris = V[F].reshape(-1, 9)
    return _orient_outward(tris, np.asarray(center, float))


def _prism(center, half_width, height, depth, shift_x=0.0):
    """Triangular prism (8 tris): apex-up triangle in x-y, extruded along z."""
    cx, cy, cz = center
    cx += shift_x
    L = (cx - half_width, cy - height / 2)
    R = (cx + half_width, cy - height / 2)
    T = (cx, cy + height / 2)
    z0, z1 = cz - depth / 2, cz + depth / 2
    p = lambda xy, z: (xy[0], xy[1], z)
    tris = [
        [p(T, z0), p(R, z0), p(L, z0)],
        [p(T, z1), p(L, z1), p(R, z1)],
        [p(L, z0), p(L, z1), p(T, z1)], [p(L, z0), p(T, z1), p(T, z0)],
        [p(T, z0), p(T, z1), p(R, z1)], [p(T, z0), p(R, z1), p(R, z0)],
        [p(R, z0), p(R, z1), p(L, z1)], [p(R, z0), p(L, z1), p(L, z0)],
    ]
    arr = np.array(tris, dtype=np.float64).reshape(-1, 9)
    return _orient_outward(arr, np.array([cx, cy - height / 6, cz]))


def _cornell_shell(b: _Builder, white, red, green):
    b.quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), white)   # floor
    b.quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), white)   # ceiling
    b.quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), white)   # back
    b.quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), red)     # left
    b.quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), green)   # right


def _box_materials():
    white = diffuse(abi.spectrum(abi.SPEC_CONST, 0.75))
    red = diffuse(abi.spectrum(abi.SPEC_STEP, 0.05, 0.8, 600.0))
    green = diffuse(abi.spectrum(abi.SPEC_GAUSS, 0.05, 0.75, 540.0, 30.0))
    light = diffuse(abi.spectrum(abi.SPEC_CONST, 0.0))
    return [white, red, green, light]


def _std_camera(res):
    return camera((0.5, 0.5, 2.7), (0.5, 0.5, 0.5), (0, 1, 0), 40.0, res, res)


def cornell_c0(res: int = 256) -> SceneData:
    b = _Builder()
    mats = _box_materials()
    _cornell_shell(b, 0, 1, 2)
    b.mesh(_box_tris((0.70, 0.66), (0.26, 0.26), 0.26, -17.0), 0)
    b.mesh(_box_tris((0.32, 0.36), (0.26, 0.26), 0.52, 20.0), 0)
    lx0, lx1, lz0, lz1, ly = 0.375, 0.625, 0.375, 0.625, 0.999
    b.quad((lx0, ly, lz0), (lx1, ly, lz0), (lx1, ly, lz1), (lx0, ly, lz1), 3)
    v, m = b.arrays()
    spd = abi.spectrum(abi.SPEC_BLACKBODY, 6500.0, 1.0, band=(380.0, 780.0))
    em = quad_emitter((lx0, ly, lz0), (lx1 - lx0, 0, 0), (0, 0, lz1 - lz0), (0, -1, 0), spd)
    return SceneData("c0_cornell", v, m, mats, [em], _std_camera(res),
                     n_photons=1 << 22, max_depth=8, rr_depth=3, hero_c=4)


def caustics_c1(res: int = 512, frame: int | None = None) -> SceneData:
    """c1; with `frame` set, the c3 animation frame of it."""
    b = _Builder()
    mats = _box_materials()
    mats.append(dielectric(abi.spectrum(abi.SPEC_SELLMEIER, *BK7_SELLMEIER)))    # 4
    mats.append(dielectric(abi.spectrum(abi.SPEC_SELLMEIER, *SF10_SELLMEIER)))   # 5
    _cornell_shell(b, 0, 1, 2)
    b.mesh(_box_tris((0.70, 0.66), (0.26, 0.26), 0.26, -17.0), 0)
    b.mesh(_box_tris((0.30, 0.34), (0.24, 0.24), 0.36, 20.0), 0)
    rot = 0.0
    shift = 0.0
    if frame is not None:
        ang = 2.0 * math.pi * frame / 64.0
        rot = ang
        shift = 0.1 * math.sin(ang)
    b.mesh(_uv_sphere((0.33, 0.62, 0.55), 0.14, 64, 64, rot_y=rot), 4)
    b.mesh(_prism((0.68, 0.60, 0.45), 0.11, 0.17, 0.22, shift_x=shift), 5)
    h = 0.02
    lc = (0.5, 0.97, 0.5)
    b.quad((lc[0] - h, lc[1], lc[2] - h), (lc[0] + h, lc[1], lc[2] - h),
           (lc[0] + h, lc[1], lc[2] + h), (lc[0] - h, lc[1], lc[2] + h), 3)
    v, m = b.arrays()
    spd = abi.spectrum(abi.SPEC_BLACKBODY, 6500.0, 1.0, band=(380.0, 780.0))
    em = quad_emitter((lc[0] - h, lc[1], lc[2] - h), (2 * h, 0, 0), (0, 0, 2 * h), (0, -1, 0), spd)
    name = "c1_caustics" if frame is None else f"c3_frame{frame:02d}"
    return SceneData(name, v, m, mats, [em], _std_camera(res),
                     n_photons=1 << 26, max_depth=16, rr_depth=0, hero_c=4)


# ---------------------------------------------------------------- icosphere + noise
def _hash3(ix, iy, iz, seed):
    """Integer lattice hash -> uniform in [-1, 1] (SplitMix64-style finalizer)."""
    x = (ix.astype(np.uint64) * np.uint64(0x9E3779B185EBCA87)) ^ \
        (iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)) ^ \
        (iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)) ^ np.uint64(seed)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def value_noise3(p: np.ndarray, seed: int, octaves: int = 3) -> np.ndarray:
    """Seeded lattice value noise, smoothstep-interpolated, `octaves` octaves with
    frequency x2 and amplitude x0.5, normalised to [-1, 1]."""
    total = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, 1.0, 0.0
    with np.errstate(over="ignore"):
        for o in range(octaves):
            q = p * freq * 4.0
            i0 = np.floor(q).astype(np.int64)
            f = q - i0
            w = f * f * (3.0 - 2.0 * f)
            acc = np.zeros(p.shape[0])
            for dx in (0, 1):
                for dy in (0, 1):
                    for dz in (0, 1):
                        h = _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed + 1013 * o)
                        wx = w[:, 0] if dx else 1 - w[:, 0]
                        wy = w[:, 1] if dy else 1 - w[:, 1]
                        wz = w[:, 2] if dz else 1 - w[:, 2]
                        acc += h * wx * wy * wz
            total += amp * acc
            norm += amp
            amp *= 0.5
            freq *= 2.0
    return total / norm


def icosphere(level: int):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 dtype=np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        key = e[:, 0] * (len(V) + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a = uk // (len(V) + 1)
        bb = uk % (len(V) + 1)
        mid = V[a] + V[bb]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        mid_idx = len(V) + inv
        nF = len(F)
        m01, m12, m20 = mid_idx[:nF], mid_idx[nF:2 * nF], mid_idx[2 * nF:]
        V = np.concatenate([V, mid], axis=0)
        F = np.concatenate([
            np.stack([F[:, 0], m01, m20], 1), np.stack([F[:, 1], m12, m01], 1),
            np.stack([F[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)], axis=0)
    return V, F


def icosphere_c2(level: int = 9, res: int = 1024, seed: int = 2024) -> SceneData:
    V, F = icosphere(level)
    disp = 0.05 * value_noise3(V, seed, octaves=3)
    radius = 0.25
    center = np.array([0.5, 0.31, 0.5])
    P = V * (radius + disp)[:, None] + center
    tris = P[F].reshape(-1, 9)
    b = _Builder()
    b.mesh(tris, 1)   # icosphere faces are wound outward already
    b.quad((-1.5, 0, -1.5), (-1.5, 0, 2.5), (2.5, 0, 2.5), (2.5, 0, -1.5), 0)
    lc, h = (0.5, 1.2, 0.5), 0.15
    b.quad((lc[0] - h, lc[1], lc[2] - h), (lc[0] + h, lc[1], lc[2] - h),
           (lc[0] + h, lc[1], lc[2] + h), (lc[0] - h, lc[1], lc[2] + h), 2)
    v, m = b.arrays()
    mats = [diffuse(abi.spectrum(abi.SPEC_CONST, 0.75)),
            dielectric(abi.spectrum(abi.SPEC_CAUCHY, 1.5, 0.004)),
            diffuse(abi.spectrum(abi.SPEC_CONST, 0.0))]
    spd = abi.spectrum(abi.SPEC_BLACKBODY, 3200.0, 1.0, band=(380.0, 780.0))
    em = quad_emitter((lc[0] - h, lc[1], lc[2] - h), (2 * h, 0, 0), (0, 0, 2 * h), (0, -1, 0), spd)
    cam = camera((0.5, 0.55, 1.9), (0.5, 0.3, 0.5), (0, 1, 0), 40.0, res, res)
    return SceneData(f"c2_icosphere_l{level}", v, m, mats, [em], cam,
                     n_photons=1 << 28, max_depth=12, rr_depth=0, hero_c=4)


def frames_c3(frame: int) -> SceneData:
    """Frame `frame` (0..63) of the batched-frames workload with its film size and
    photon budget: 16 frames each at 128^2, 256^2, 512^2, 768^2; budget
    2^20 * px / 128^2 capped at 2^24."""
    res = (128, 256, 512, 768)[frame // 16]
    s = caustics_c1(res=res, frame=frame)
    s.n_photons = min(1 << 24, (1 << 20) * (res * res) // (128 * 128))
    return s


def soup(n_tris: int, seed: int, res: int = 256) -> SceneData:
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.0, 1.0, size=(n_tris, 3))
    # Two edges from the first vertex with lengths uniform in [0.002, 0.02] in random
    # directions; the third edge follows from them.
    L1 = rng.uniform(0.002, 0.02, n_tris)
    L2 = rng.uniform(0.002, 0.02, n_tris)
    d1 = rng.normal(size=(n_tris, 3))
    d1 /= np.linalg.norm(d1, axis=1, keepdims=True)
    d2 = rng.normal(size=(n_tris, 3))
    d2 -= d1 * np.einsum("ij,ij->i", d1, d2)[:, None] * 0.5
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    v0 = centers
    v1 = v0 + d1 * L1[:, None]
    v2 = v0 + d2 * L2[:, None]
    tris = np.concatenate([v0, v1, v2], axis=1)
    levels = 1024
    albedo_idx = rng.integers(0, levels, n_tris)
    mats = [diffuse(abi.spectrum(abi.SPEC_CONST, 0.2 + 0.7 * (k + 0.5) / levels)) for k in range(levels)]
    mats.append(diffuse(abi.spectrum(abi.SPEC_CONST, 0.0)))
    b = _Builder()
    b.mesh(tris, albedo_idx)
    lc, h = (0.5, 1.3, 0.5), 0.25
    b.quad((lc[0] - h, lc[1], lc[2] - h), (lc[0] + h, lc[1], lc[2] - h),
           (lc[0] + h, lc[1], lc[2] + h), (lc[0] - h, lc[1], lc[2] + h), levels)
    v, m = b.arrays()
    spd = abi.spectrum(abi.SPEC_BLACKBODY, 5000.0, 1.0, band=(380.0, 780.0))
    em = quad_emitter((lc[0] - h, lc[1], lc[2] - h), (2 * h, 0, 0), (0, 0, 2 * h), (0, -1, 0), spd)
    cam = camera((0.5, 0.5, 2.6), (0.5, 0.5, 0.5), (0, 1, 0), 40.0, res, res)
    return SceneData(f"soup_{n_tris}_{seed}", v, m, mats, [em], cam,
                     n_photons=1 << 22, max_depth=8, rr_depth=0, hero_c=4)


def glossy_box(res: int = 128) -> SceneData:
    """Cornell box whose tall box is a rough conductor (Glossy, roughness 0.3); a
    parity case for the conductor lobe (the configs themselves have none)."""
    s = cornell_c0(res)
    mats = list(s.materials)
    mats.append(glossy(abi.spectrum(abi.SPEC_GAUSS, 0.3, 0.6, 620.0, 60.0), 0.3))
    m = s.mat_ids.copy()
    m[22:34] = len(mats) - 1
    return SceneData("glossy_box", s.verts, m, mats, list(s.emitters), s.camera,
                     n_photons=1 << 20, max_depth=8, rr_depth=0, hero_c=4)


def mirror_box(res: int = 128) -> SceneData:
    """Cornell box with a tinted mirror panel hung 1 mm in front of the (red) left wall
    and the short box turned into a mirror (six rotated planes, two triangles each); a
    parity case for the flat-mirror camera connection (the configs themselves have
    no mirrors). The camera sees the right half of the room in the panel."""
    s = cornell_c0(res)
    mats = list(s.materials)
    mats.append(mirror(abi.spectrum(abi.SPEC_STEP, 0.55, 0.92, 520.0)))
    mi = len(mats) - 1
    m = s.mat_ids.copy()
    m[10:22] = mi
    x = 0.001
    panel = np.array([[x, 0.08, 0.1, x, 0.08, 0.92, x, 0.82, 0.92],
                      [x, 0.08, 0.1, x, 0.82, 0.92, x, 0.82, 0.1]], np.float32)
    v = np.concatenate([s.verts, panel], axis=0)
    m = np.concatenate([m, np.array([mi, mi], np.int32)])
    return SceneData("mirror_box", v, m, mats, list(s.emitters), s.camera,
                     n_photons=1 << 20, max_depth=8, rr_depth=0, hero_c=4)


CONFIGS = {
    "c0": cornell_c0,
    "c1": caustics_c1,
    "c2": icosphere_c2,
}
