"""Film -> image on the CPU side: the oracle's restatement of filmToLinear /
filmToRgb8 (src/main.cpp:10282-10348, src/color.h:41-53) pinned by what the
mathematics and the sRGB standard fix, and the PFM / PPM / PNG writers.

The GPU reduction (gp_film_to_linear, gp_film_tonemap) is compared with this oracle
bit for bit in test_gpu_image.py."""
import struct
import zlib

import numpy as np
import pytest

import goodphotons_amd as gp

# IEC 61966-2-1 linear sRGB -> XYZ (D65). The reference's matrix is its inverse to 4
# decimals, so each primary must come back as a unit vector to ~1e-3.
SRGB_TO_XYZ = np.array([[0.4124, 0.3576, 0.1805],
                        [0.2126, 0.7152, 0.0722],
                        [0.0193, 0.1192, 0.9505]])


def _film_for_linear(orc, lin_target, n_photons):
    """XYZ film sums whose linear sRGB is lin_target (approximately, via the standard matrix)."""
    xyz = np.asarray(lin_target, float) @ SRGB_TO_XYZ.T
    return xyz * n_photons * orc.cie_y_integral()


def test_primaries_and_white(orc):
    n = 1000.0
    lin = orc.film_to_linear(_film_for_linear(orc, np.eye(3), n), n)
    assert np.allclose(lin, np.eye(3), atol=2e-3)
    d65 = np.array([[0.95047, 1.0, 1.08883]]) * n * orc.cie_y_integral()
    assert np.allclose(orc.film_to_linear(d65, n), 1.0, atol=3e-4)


def test_linear_in_film_and_inverse_in_photons(orc):
    rng = np.random.default_rng(3)
    xyz = rng.random((50, 3))
    a = orc.film_to_linear(xyz, 10.0)
    assert np.allclose(orc.film_to_linear(2.0 * xyz, 10.0), 2.0 * a, rtol=1e-14, atol=1e-18)
    assert np.allclose(orc.film_to_linear(xyz, 20.0), 0.5 * a, rtol=1e-14, atol=1e-18)


def test_srgb_encoding_textbook_values(orc):
    # absolute EV: exposure = 6 x comp; pick comp so exposure = 1 and the linear values
    # land on known sRGB codes: 0 -> 0, 0.18 -> 118, 0.5 -> 188, 1 -> 255, 2 -> 255.
    vals = np.array([0.0, 0.18, 0.5, 1.0, 2.0, 0.0031308])
    lin = np.repeat(vals[:, None], 3, axis=1)
    n = 1.0
    xyz = _film_for_linear(orc, lin, n)
    # undo the 4-decimal matrix mismatch exactly: feed the oracle's own inverse
    m_inv = np.linalg.inv(orc.film_to_linear(np.eye(3) * n * orc.cie_y_integral(), n).T)
    xyz = (lin @ m_inv.T) * n * orc.cie_y_integral()
    img, e_auto, exposure, _ = orc.film_to_rgb8(xyz, len(vals), 1, n, exp_comp=1.0 / 6.0, absolute=True)
    assert e_auto == 6.0 and abs(exposure - 1.0) < 1e-15
    codes = img[0, :, 0]
    assert list(codes[:5]) == [0, 118, 188, 255, 255]
    assert codes[5] == int(12.92 * 0.0031308 * 255 + 0.5)      # linear segment: 10


def test_p99_anchor_is_the_order_statistic(orc):
    rng = np.random.default_rng(11)
    w, h = 37, 23
    xyz = rng.gamma(0.3, 1.0, (w * h, 3))
    xyz[rng.random(w * h) < 0.2] = 0.0
    n = 5000.0
    lin = orc.film_to_linear(xyz, n)
    lum = np.maximum(lin.max(axis=1), 0.0)
    k = int(0.99 * (w * h - 1))
    p99 = np.sort(lum)[k]
    img, e_auto, exposure, anchor = orc.film_to_rgb8(xyz, w, h, n, anchor=0.0)
    assert e_auto == 0.9 / p99 and exposure == e_auto and anchor == e_auto
    # exposure compensation multiplies; a locked anchor is reused as is
    _, e2, x2, _ = orc.film_to_rgb8(xyz, w, h, n, exp_comp=2.0)
    assert e2 == e_auto and x2 == 2.0 * e_auto
    _, e3, x3, a3 = orc.film_to_rgb8(xyz * 7.0, w, h, n, anchor=anchor)
    assert e3 == anchor and a3 == anchor
    # an all-dark film falls back to exposure 1
    _, e4, _, _ = orc.film_to_rgb8(np.zeros((w * h, 3)), w, h, n)
    assert e4 == 1.0


def test_rows_flip_to_image_top(orc):
    w, h = 8, 5
    xyz = np.zeros((w * h, 3))
    x0, y0 = 3, 1                       # film row 1 counts from the bottom
    xyz[y0 * w + x0] = [1.0, 1.0, 1.0]
    img, _, _, _ = orc.film_to_rgb8(xyz, w, h, 1.0, absolute=True, exp_comp=1e6)
    ys, xs = np.nonzero(img.max(axis=2))
    assert list(ys) == [h - 1 - y0] and list(xs) == [x0]


def test_image_writers(tmp_path):
    rng = np.random.default_rng(5)
    lin = rng.random((4, 6, 3)).astype(np.float32)
    p = tmp_path / "a.pfm"
    gp.write_pfm(p, lin)
    raw = p.read_bytes()
    head = b"PF\n6 4\n-1.0\n"
    assert raw.startswith(head)
    assert np.array_equal(np.frombuffer(raw[len(head):], "<f4").reshape(4, 6, 3), lin)
    img = rng.integers(0, 256, (4, 6, 3), dtype=np.uint8)
    q = tmp_path / "a.ppm"
    gp.write_ppm(q, img)
    assert q.read_bytes() == b"P6\n6 4\n255\n" + img.tobytes()
    r = tmp_path / "a.png"
    gp.write_png(r, img)
    data = r.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    off, chunks = 8, {}
    while off < len(data):
        ln = struct.unpack(">I", data[off:off + 4])[0]
        t = data[off + 4:off + 8]
        chunks[t] = chunks.get(t, b"") + data[off + 8:off + 8 + ln]
        assert struct.unpack(">I", data[off + 8 + ln:off + 12 + ln])[0] == zlib.crc32(t + data[off + 8:off + 8 + ln])
        off += 12 + ln
    assert struct.unpack(">II", chunks[b"IHDR"][:8]) == (6, 4)
    rows = zlib.decompress(chunks[b"IDAT"])
    dec = np.frombuffer(rows, np.uint8).reshape(4, 1 + 18)
    assert (dec[:, 0] == 0).all() and np.array_equal(dec[:, 1:].reshape(4, 6, 3), img)
