"""Device film checkpoints (.ftbuf, src/main.cpp:12443-12483): the file is an FTBUF01
sidecar the reference's readCheckpoint accepts (checked with the oracle's restatement
of the reader), and a render saved, reloaded and continued is bit-identical to one
that never stopped and to a single-shot render of all photons."""
import struct

import numpy as np
import pytest

import goodphotons_amd as gp
from goodphotons_amd import scenes

pytestmark = pytest.mark.gpu


def test_save_resume_is_bit_exact(tmp_path, orc, gplib):
    sd = scenes.caustics_c1(96)
    gs = gp.GpuScene(sd, gplib)
    n1, n2 = 1 << 17, 3 << 15
    kw = dict(hero_c=4, max_depth=sd.max_depth)
    guard = gplib.lib.gp_checkpoint_guard(int(sd.n_tris), 0, len(sd.emitters), 1.25, ord("B"), 96, 96)
    # continuous: one film, two traces
    ref = gp.Film(96, 96, max_photons=n1 + n2, lib=gplib)
    e1, _ = gp.trace(gs, sd.camera, ref, n1, seed_base=0, **kw)
    path = tmp_path / "c1.png.ftbuf"
    ref.save_checkpoint(path, n1, energy=e1, guard=guard)
    gp.trace(gs, sd.camera, ref, n2, seed_base=n1, **kw)
    # the file is what readCheckpoint reads
    rc, N, en, xyz, hits, trailing = orc.read_checkpoint(path, 96, 96, guard)
    assert rc == 0 and N == n1 and trailing > 0
    assert np.allclose(en, [e1.emitted, e1.absorbed, e1.sensor, e1.escaped, e1.residual], rtol=0, atol=0)
    assert orc.read_checkpoint(path, 96, 96, guard + 1)[0] == 2
    assert orc.read_checkpoint(path, 96, 97, guard)[0] == 2
    # resume and continue
    film, n_done, e_loaded = gp.Film.load_checkpoint(path, guard=guard, lib=gplib)
    assert n_done == n1 and e_loaded.emitted == e1.emitted
    x0, h0 = film.download()
    assert np.array_equal(x0, xyz) and np.array_equal(h0, hits)
    gp.trace(gs, sd.camera, film, n2, seed_base=n_done, **kw)
    w_res, s_res = film.raw()
    w_ref, s_ref = ref.raw()
    assert s_res == s_ref and np.array_equal(w_res, w_ref)
    # single shot of all photons
    one = gp.Film(96, 96, max_photons=n1 + n2, lib=gplib)
    gp.trace(gs, sd.camera, one, n1 + n2, seed_base=0, **kw)
    assert np.array_equal(one.raw()[0], w_ref)


def test_load_refuses_mismatch_and_reference_only_files(tmp_path, gplib):
    sd = scenes.cornell_c0(32)
    gs = gp.GpuScene(sd, gplib)
    film = gp.Film(32, 32, max_photons=1 << 14, lib=gplib)
    e, _ = gp.trace(gs, sd.camera, film, 1 << 14, hero_c=4, max_depth=8)
    path = tmp_path / "c0.ftbuf"
    film.save_checkpoint(path, 1 << 14, energy=e, guard=77)
    with pytest.raises(gp.GpError) as ei:
        gp.Film.load_checkpoint(path, guard=78, lib=gplib)
    assert ei.value.status == 8          # GP_ERR_STATE
    # strip the fixed-point section: what the reference itself writes
    raw = path.read_bytes()
    head = 8 + 4 * 3 + 8 + 5 * 8 + 8 + 32 * 32 * 4 * 8
    ref_only = tmp_path / "ref.ftbuf"
    ref_only.write_bytes(raw[:head])
    with pytest.raises(gp.GpError) as ei:
        gp.Film.load_checkpoint(ref_only, guard=77, lib=gplib)
    assert ei.value.status == 2
    # a fixed-point section with a scale that is not positive, or a saturation value
    # that is not a power of two, is refused rather than loaded
    for off, val in ((8, struct.pack("<d", -1.0)), (16, struct.pack("<Q", 3))):
        bad = bytearray(raw)
        bad[head + off:head + off + 8] = val
        p = tmp_path / f"bad{off}.ftbuf"
        p.write_bytes(bytes(bad))
        with pytest.raises(gp.GpError) as ei:
            gp.Film.load_checkpoint(p, guard=77, lib=gplib)
        assert ei.value.status == 1      # GP_ERR_INVALID
