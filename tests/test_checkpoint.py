"""Checkpoint identity guard (checkpointGuard, src/main.cpp:12317-12329), CPU side:
the oracle's restatement and the library's host function agree, and both reduce
to the published FNV-1a 64-bit test vector for a single mixed byte."""
import numpy as np

import goodphotons_amd as gp

FNV_PRIME = 1099511628211
FNV1A_OF_A = 0xAF63DC4C8601EC8C      # FNV-1a 64 of the one-byte string "a"


def test_guard_reduces_to_fnv1a_test_vector(orc):
    # one mix of 0x61 from the offset basis is FNV-1a("a"); six zero words follow
    want = (FNV1A_OF_A * FNV_PRIME ** 6) % (1 << 64)
    assert orc.checkpoint_guard(0x61, 0, 0, 0.0, 0, 0, 0) == want
    assert gp.load().lib.gp_checkpoint_guard(0x61, 0, 0, 0.0, 0, 0, 0) == want


def test_guard_library_equals_oracle(orc):
    lib = gp.load().lib
    rng = np.random.default_rng(9)
    for _ in range(50):
        a = [int(x) for x in rng.integers(0, 1 << 40, 3)]
        p = float(rng.random() * 10.0 ** int(rng.integers(-5, 25)))
        rx, ry = [int(x) for x in rng.integers(1, 4096, 2)]
        m = int(rng.choice([ord("A"), ord("B"), ord("C")]))
        assert lib.gp_checkpoint_guard(a[0], a[1], a[2], p, m, rx, ry) == orc.checkpoint_guard(a[0], a[1], a[2], p, m, rx, ry)
    # every field changes the guard
    base = (36, 0, 1, 3.5, ord("B"), 256, 256)
    g0 = orc.checkpoint_guard(*base)
    for k in range(7):
        v = list(base)
        v[k] = v[k] + (1 if k != 3 else 1e-9)
        assert orc.checkpoint_guard(*v) != g0
