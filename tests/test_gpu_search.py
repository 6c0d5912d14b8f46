"""SearchNeighbors on the GPU against the oracle: bit-exact pair lists.

Both sides emit the pairs of each fluid query in (d2, j) rank order, so the
(i_pair, j_pair, rij2) arrays must be identical element by element, rij2
compared as FP64 bit patterns.
"""
import numpy as np
import pytest

from conftest import load_golden_scene
from uph_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-search] {msg}", flush=True)


def _check(sc, name, K=40):
    from oracle import oracle as O
    from uph_amd import solver
    _say(f"{name}: n={sc.n} nf={sc.n_fluid}")
    ref = O.search(sc.pos, sc.n_fluid, sc.n, sc.particle_boundary, K, sc.cutoff2, nthreads=0)
    got = solver.search_neighbors(sc.pos, sc.n_fluid, sc.n, sc.particle_boundary, K, sc.cutoff2)
    assert len(got[0]) == len(ref[0]), (len(got[0]), len(ref[0]))
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1])
    assert np.array_equal(got[2].view(np.uint64), ref[2].view(np.uint64))
    return got


def test_dam_break_5k_lattice():
    _check(scenes.dam_break(), "dam break 5K lattice")


def test_jittered_lattice():
    _check(scenes.dam_break(jitter=0.05, seed=11), "dam break 5K jittered +-0.05dx")


def test_reference_hydrostatic_deck():
    _check(load_golden_scene("hydrostatic_init.npz"), "tests/hydrostaticPressure FluidParticleInit")


def test_reference_dambreak_deck_cap_binds():
    _check(load_golden_scene("dambreak.npz"), "tests/dambreak (K=40 binds)")


def test_reference_waves_deck():
    _check(load_golden_scene("waves.npz"), "tests/waves")


def test_small_K():
    _check(scenes.dam_break(width=0.3, height=0.3, jitter=0.1, seed=3), "K=5", K=5)


def test_row_table_signature_matches():
    from uph_amd import solver
    sc = scenes.dam_break(width=0.4, height=0.4, jitter=0.05, seed=5)
    _say("double** SearchNeighbors")
    a = solver.search_neighbors(sc.pos, sc.n_fluid, sc.n, sc.particle_boundary, 40, sc.cutoff2)
    cap = sc.n_fluid * 40
    ip = np.zeros(cap, np.int32); jp = np.zeros(cap, np.int32); r2 = np.zeros(cap)
    m = solver.search_neighbors_rows(sc.pos, ip, jp, r2, sc.n_fluid, sc.n,
                                     sc.particle_boundary, 40, sc.cutoff2)
    assert m == len(a[0])
    assert np.array_equal(ip[:m], a[0]) and np.array_equal(jp[:m], a[1])
    assert np.array_equal(r2[:m], a[2])


def test_edge_cases():
    from uph_amd import solver
    pos = np.array([[0.0, 0.0], [0.5, 0.0], [0.0, 0.5]])
    _say("edge cases")
    assert len(solver.search_neighbors(pos, 0, 3, 2, 40, 1.0)[0]) == 0
    ip, jp, _ = solver.search_neighbors(pos, 3, 3, 2, 40, 1.0)
    assert list(zip(ip, jp)) == [(0, 1)]
    ip, jp, _ = solver.search_neighbors(pos, 3, 3, 3, 1, 1.0)
    assert list(zip(ip, jp)) == [(0, 1)]
    assert len(solver.search_neighbors(pos, 1, 3, 3, 40, 0.25)[0]) == 0
    with pytest.raises(RuntimeError):
        solver.search_neighbors(pos, 3, 3, 3, 40, 1.0, capacity=1)
    # widely spread particles: the grid is coarsened, the result must not change
    rng = np.random.default_rng(4)
    far = np.concatenate([rng.uniform(0, 0.05, (300, 2)), [[1e4, -1e4], [-3e3, 2e3]]])
    ref_sc = scenes.Scene(pos=far, vel=np.zeros_like(far), rho=np.full(len(far), 1000.0),
                          pressure=np.zeros(len(far)), mass=np.ones(len(far)),
                          itype=np.ones(len(far), np.int32), lvel=np.zeros_like(far),
                          n_fluid=len(far), n_fixed=0, n_moving=0, dx=0.002, h=0.004)
    _check(ref_sc, "spread-out positions (coarsened grid)")


def test_dam_break_1m_lattice():
    sc = scenes.dam_break(dx=2 ** 0.5 * 1e-3)
    got = _check(sc, "dam break 1M lattice")
    # 20 neighbours per interior particle: 10 pairs each
    assert len(got[0]) > 9.5 * sc.n_fluid
