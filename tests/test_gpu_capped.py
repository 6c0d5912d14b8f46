"""The capped (K-nearest) path's hand-off from the per-lane K-th search to the capped
pass, bit for bit: the per-lane search leaves each capped query's pair partners as
bits (kmask: the admitted candidates above it by original index, and the candidates
below it, which wait for their own K-th keys), and the capped pass walks those
without testing distances again.  Four instances are checked against runs that take
other code paths over the same inputs:

* a batch holding a capped scene without mirror sections (the reference's dambreak
  deck, virtualType 77, h = 2.691 dx: about 88 particles in range, K = 40) and an
  uncapped one, against each scene stepped alone (the multi-scene instance of the
  per-lane kernel against the one-scene search);
* a batch of two scenes whose queries are all capped (the K-th key search alone first,
  with the scene per lane) against each scene alone (its candidate rows staged in LDS);
* the reference scheme with PolyOrderSeven on that deck, the K-th key search and the
  capped bit walk with the masks against the direct kernel (no mask);
* hipGraph replay of capped steps against eager launches, with eager work and
  synchronisations between the replays.
"""
import numpy as np
import pytest

from conftest import load_golden_scene
from uph_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-capped] {msg}", flush=True)


def _moving(sc, seed=3, amp=0.1):
    rng = np.random.default_rng(seed)
    nf = sc.n_fluid
    sc.vel[:nf] = rng.normal(0, amp, (nf, 2))
    sc.lvel[:nf] = sc.vel[:nf] + rng.normal(0, 0.1 * amp, (nf, 2))
    return sc


def _same(a, b, what):
    for key in ("x", "v", "rho"):
        assert np.array_equal(a[key], b[key]), (what, key)


def test_batch_with_capped_scene_equals_single_scenes():
    from uph_amd.solver import GpuSolver
    capped = _moving(load_golden_scene("dambreak.npz"))
    plain = scenes.dam_break(width=0.4, height=0.5, dx=0.02, jitter=0.05, seed=2)
    _say(f"batch: capped deck ({capped.n} rows) + dam break ({plain.n} rows), 15 steps")
    gb = GpuSolver([capped, plain])
    gb.step(15)
    assert gb.sync() == 15            # the capped pass ran in every step
    for k, sc in enumerate((capped, plain)):
        gs = GpuSolver(sc, sort_every_step=True)   # as the batch, which holds a capped scene
        gs.step(15)
        gs.sync()
        _same(gb.state(k), gs.state(0), f"scene {k}")


def test_batch_of_capped_scenes_equals_single_scenes():
    """Every scene expects more than K in range and none has mirror sections: the batch
    takes the K-th key search alone first with the scene per lane (no staged rows), each
    scene alone the one-scene instance (candidate rows staged in LDS), and the capped
    pass computes every slot in both."""
    from uph_amd.solver import GpuSolver
    deck = _moving(load_golden_scene("dambreak.npz"))
    gen = _moving(scenes.dam_break(width=0.3, height=0.4, dx=0.01, tank=(0.8, 0.8), jitter=0.05, seed=6,
                                   h_factor=2.691),
                  seed=4)
    _say(f"batch: capped deck ({deck.n} rows) + capped dam break ({gen.n} rows), 12 steps")
    gb = GpuSolver([deck, gen])
    gb.step(12)
    assert gb.sync() == 12
    for k, sc in enumerate((deck, gen)):
        gs = GpuSolver(sc)
        gs.step(12)
        assert gs.sync() == 12
        _same(gb.state(k), gs.state(0), f"scene {k}")


def test_ref_poly7_capped_compact_equals_direct():
    from uph_amd.solver import GpuSolver
    sc = load_golden_scene("dambreak.npz", mode=scenes.MODE_REF, kernel=scenes.KERNEL_POLY7, alpha_av=0.0)
    sc = _moving(sc)
    rng = np.random.default_rng(5)
    sc.pressure[:sc.n_fluid] = np.abs(rng.normal(500.0, 200.0, sc.n_fluid))
    _say("ref mode, PolyOrderSeven, tests/dambreak: compacted vs direct, 12 steps")
    a = GpuSolver(sc, pair_variant=0)
    b = GpuSolver(sc, pair_variant=1)
    a.step(12)
    b.step(12)
    assert a.sync() == 12 and b.sync() == 12
    _same(a.state(), b.state(), "ref poly7")
    assert np.array_equal(a.pressure(), b.pressure())


def test_capped_graph_replay_equals_eager():
    from uph_amd.solver import GpuSolver
    sc = _moving(load_golden_scene("dambreak.npz"), seed=8)
    _say("capped deck: hipGraph replay vs eager")
    ge = GpuSolver(sc)
    gg = GpuSolver(sc, use_graph=True)
    done = 0
    for chunk in (1, 4, 10):
        ge.step(chunk)
        gg.step(chunk)
        done += chunk
        assert ge.sync() == done and gg.sync() == done
        _same(ge.state(), gg.state(), f"after {done} steps")
