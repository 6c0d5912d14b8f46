"""uph_search_neighbors_noalloc (the search with every device buffer in one fixed
file-scope array) against uph_search_neighbors_cap (the search with its own
allocations), through the C ABI: two calls in a row, the second on a smaller input,
each compared bit for bit with the allocating entry point on the same positions."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_golden_scene
from uph_amd import scenes

pytestmark = pytest.mark.gpu

MAX_PARTICLES = 65536


def _say(msg):
    print(f"[no-alloc] {msg}", flush=True)


def _noalloc(pos, nf, n, pb, K, dist, cap):
    from uph_amd import solver
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 2)
    ip = np.full(cap, -7, np.int32)
    jp = np.full(cap, -7, np.int32)
    r2 = np.full(cap, -7.0)
    m = solver.lib().uph_search_neighbors_noalloc(
        pos.ctypes.data_as(C.POINTER(C.c_double)), ip.ctypes.data_as(C.POINTER(C.c_int)),
        jp.ctypes.data_as(C.POINTER(C.c_int)), r2.ctypes.data_as(C.POINTER(C.c_double)),
        int(nf), int(n), int(pb), int(K), float(dist), int(cap))
    return m, ip, jp, r2


def _check(sc, name):
    from uph_amd import solver
    args = (sc.n_fluid, sc.n, sc.particle_boundary, sc.K, sc.cutoff2)
    ref = solver.search_neighbors(sc.pos, *args)
    cap = max(1, sc.n_fluid * sc.K)
    m, ip, jp, r2 = _noalloc(sc.pos, *args, cap)
    _say(f"{name}: n={sc.n}, {m} pairs (allocating search: {len(ref[0])})")
    assert m == len(ref[0]), solver.lib().uph_search_last_error()
    assert np.array_equal(ip[:m], ref[0])
    assert np.array_equal(jp[:m], ref[1])
    assert np.array_equal(r2[:m], ref[2])
    assert np.all(ip[m:] == -7)            # nothing written past the count


def test_noalloc_search_equals_allocating_search_twice():
    # first call: the reference's dambreak deck (h = 2.691 dx, K = 40 binds: the
    # capped K-th key search runs); second call: a smaller jittered dam break
    big = load_golden_scene("dambreak.npz")
    small = scenes.dam_break(width=0.3, height=0.4, dx=0.02, jitter=0.05, seed=3)
    assert small.n < big.n <= MAX_PARTICLES
    _check(big, "first call, tests/dambreak deck")
    _check(small, "second call, smaller dam break")


def test_noalloc_search_limits_are_errors():
    from uph_amd import solver
    sc = scenes.dam_break(width=0.3, height=0.4, dx=0.02)
    args = (sc.n_fluid, sc.n, sc.particle_boundary, sc.K, sc.cutoff2)
    m, ip, *_ = _noalloc(sc.pos, *args, 5)       # capacity below the pair count
    assert m == -2 and np.all(ip == -7)
    assert b"capacity" in solver.lib().uph_search_last_error()
    n = MAX_PARTICLES + 1                        # rejected before any device work
    pos = np.random.default_rng(0).uniform(0.0, 1.0, (n, 2))
    m, ip, *_ = _noalloc(pos, n, n, n - 1, 40, 1e-4, 10)
    assert m == -3 and np.all(ip == -7)
    assert b"maximum" in solver.lib().uph_search_last_error()
    # the allocating entry point's own message is reported after its own calls
    ref = solver.search_neighbors(sc.pos, *args)
    assert len(ref[0]) > 0 and solver.lib().uph_search_last_error() == b""
    _check(sc, "after the rejected calls")
