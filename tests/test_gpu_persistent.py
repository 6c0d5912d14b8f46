"""uph_step_persistent (all phases of n steps in one cooperative launch, grid-wide barriers
between the phases) against uph_step (one launch per kernel), bit for bit, on handles
that re-sort every step (the persistent step's phases)."""
import numpy as np
import pytest

from uph_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-persistent] {msg}", flush=True)


def _same(a, b, what):
    for key in ("x", "v", "rho"):
        assert np.array_equal(a[key], b[key]), (what, key)


def test_persistent_batch_equals_stepped():
    """Three scenes of different sizes (wavefronts holding two scenes' slots included)."""
    from uph_amd.solver import GpuSolver
    scs = [scenes.dam_break(width=0.4, height=0.5, dx=0.02),
           scenes.dam_break(width=0.3, height=0.6, dx=0.015),
           scenes.dam_break(width=0.5, height=0.3, dx=0.025, jitter=0.05, seed=1)]
    _say("batch of 3: persistent vs stepped, 7 + 23 steps")
    a, b = GpuSolver(scs, sort_every_step=True), GpuSolver(scs, sort_every_step=True)
    for chunk in (7, 23):
        a.step_persistent(chunk)
        b.step(chunk)
        assert a.sync() == b.sync() == 0
        for k in range(len(scs)):
            _same(a.state(k), b.state(k), f"scene {k}")


def test_persistent_capped_steps_equal_stepped():
    """One scene with a patch compressed to 0.7 dx: its queries have more than K = 29 in
    range, so the capped phase runs in every step, while the handle (20 in range on the
    lattice) is not one whose queries are expected to be capped."""
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(width=0.5, height=0.6, dx=0.02)
    sc.K = 29
    nf = sc.n_fluid
    x = sc.pos[:nf]
    patch = (x[:, 0] < 0.16) & (x[:, 1] < 0.16)
    sc.pos[:nf][patch] = x[patch] * 0.7
    _say(f"capped patch of {int(patch.sum())} particles: persistent vs stepped, 12 steps")
    a, b = GpuSolver(sc, sort_every_step=True), GpuSolver(sc, sort_every_step=True)
    a.step_persistent(12)
    b.step(12)
    ca, cb = a.sync(), b.sync()
    assert ca == cb == 12
    _same(a.state(), b.state(), "capped")


def test_persistent_rejects_unsupported_handles():
    from uph_amd.solver import GpuSolver
    capped = scenes.dam_break(width=0.3, height=0.3, dx=0.02, h_factor=2.0)
    with pytest.raises(RuntimeError, match="persistent"):
        GpuSolver(capped).step_persistent(1)
    mirrored = scenes.dam_break(width=0.3, height=0.3, dx=0.02)
    mirrored.vtype, mirrored.flat_bottom, mirrored.depth = 8, 4.0, 0.3
    with pytest.raises(RuntimeError, match="persistent"):
        GpuSolver(mirrored).step_persistent(1)
    with pytest.raises(RuntimeError, match="persistent"):
        GpuSolver(scenes.dam_break(width=0.3, height=0.3, dx=0.02), pair_variant=1,
                  sort_every_step=True).step_persistent(1)
    with pytest.raises(RuntimeError, match="persistent"):   # a lazy re-sort handle
        GpuSolver(scenes.dam_break(width=0.3, height=0.3, dx=0.02)).step_persistent(1)
