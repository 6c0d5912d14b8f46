"""Lazy re-sort (the default for handles without mirror rows or capped queries): cells
grow by a skin, the particles keep their sorted slots until one of a scene's particles
has left its binned cell by more than the growth, the pair kernel's epilogue predicts
the next step's neighbour mirrors, and the step's two state sets exchange roles.

Pair sets are the exact in-range sets either way; what changes is the order of the
slots, so the sums round differently from a handle that sorts every step.  The tests
check the path against the FP64 oracle, against a sort-every-step handle, and for the
invariances the design claims: results do not depend on how the steps are split into
calls (the last step of a call holds its re-sort over to the next call), a scene in a
batch equals its single-scene run, and state set from outside restarts the sort."""
import numpy as np
import pytest

from uph_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-lazy] {msg}", flush=True)


def _same(a, b, what):
    for key in ("x", "v", "rho"):
        assert np.array_equal(a[key], b[key]), (what, key)


def _rel(a, b, n):
    return float(np.max(np.abs(a["rho"][:n] - b["rho"][:n]) / b["rho"][:n]))


def _dx_err(a, b, n, dx):
    return float(np.max(np.hypot(*(a["x"][:n] - b["x"][:n]).T)) / dx)


def _scene(seed=2, jitter=0.1):
    return scenes.dam_break(width=0.5, height=0.5, dx=0.02, jitter=jitter, seed=seed)


def test_lazy_matches_oracle_with_resorts():
    """A small skin (2 % of the cell) makes the jittered dam break re-sort every few
    steps; 300 steps stay within the oracle's tolerance."""
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    sc = _scene()
    g = GpuSolver(sc, skin=0.02)
    g.step(300)
    assert g.sync() == 0
    nres = g.resorts()
    o = Sim(sc, nthreads=0)
    o.step(300)
    n = sc.n_fluid
    drho, dxe = _rel(g.state(), o.state(), n), _dx_err(g.state(), o.state(), n, sc.dx)
    _say(f"skin 0.02: {nres} re-sorts in 300 steps; vs oracle rho {drho:.2e}, x {dxe:.2e} dx")
    assert 1 < nres <= 301
    assert drho < 1e-6 and dxe < 1e-4


def test_lazy_default_skin_resorts_rarely_and_matches_sorting_every_step():
    """The default skin (2 % for one scene): fewer re-sorts than steps; against a handle
    that sorts every step the results differ only by the summation order (rounding)."""
    from uph_amd.solver import GpuSolver
    sc = _scene()
    a = GpuSolver(sc)
    b = GpuSolver(sc, sort_every_step=True)
    a.step(300)
    b.step(300)
    assert a.sync() == b.sync() == 0
    nres = a.resorts()
    assert b.resorts() == -1 and b.skin() == 0.0 and a.skin() == 0.02
    n = sc.n_fluid
    drho, dxe = _rel(a.state(), b.state(), n), _dx_err(a.state(), b.state(), n, sc.dx)
    _say(f"default skin: {nres} re-sorts in 300 steps; vs sorting every step rho {drho:.2e}, x {dxe:.2e} dx")
    assert 1 <= nres < 150
    assert drho < 1e-8 and dxe < 1e-6


@pytest.mark.parametrize("skin", [0.02, 0.05])
def test_lazy_results_do_not_depend_on_call_chunking(skin):
    """One call of 120 steps, calls of 7 and 1 steps, and calls of 1 step with a state
    read between them: identical bit for bit (a re-sort that the last step of a call
    holds over runs at the next call's first step, between the same two steps)."""
    from uph_amd.solver import GpuSolver
    sc = _scene(seed=4)
    a, b, c = GpuSolver(sc, skin=skin), GpuSolver(sc, skin=skin), GpuSolver(sc, skin=skin)
    a.step(120)
    done = 0
    for chunk in (7, 1, 7, 1, 7, 1, 96):
        b.step(chunk)
        done += chunk
    assert done == 120
    for _ in range(120):
        c.step(1)
        c.state()
    assert a.sync() == b.sync() == c.sync() == 0
    _say(f"skin {skin}: {a.resorts()} / {b.resorts()} / {c.resorts()} re-sorts")
    _same(a.state(), b.state(), "chunks")
    _same(a.state(), c.state(), "single steps")


def test_lazy_batch_equals_single_scenes():
    """Three scenes of different sizes and speeds in one batch, with re-sorts: each
    equals its own single-scene run bit for bit (wavefronts holding two scenes' slots
    included)."""
    from uph_amd.solver import GpuSolver
    scs = [scenes.dam_break(width=0.4, height=0.5, dx=0.02, jitter=0.1, seed=1),
           scenes.dam_break(width=0.3, height=0.6, dx=0.015),
           scenes.dam_break(width=0.5, height=0.3, dx=0.025, jitter=0.05, seed=3)]
    gb = GpuSolver(scs, skin=0.03)
    gb.step(150)
    assert gb.sync() == 0
    _say(f"batch of 3: {gb.resorts()} re-sorts in 150 steps")
    for k, sc in enumerate(scs):
        gs = GpuSolver(sc, skin=0.03)
        gs.step(150)
        _same(gb.state(k), gs.state(0), f"scene {k}")


def test_lazy_compact_equals_direct():
    from uph_amd.solver import GpuSolver
    sc = _scene(seed=6)
    a, b = GpuSolver(sc, skin=0.02), GpuSolver(sc, skin=0.02, pair_variant=1)
    a.step(80)
    b.step(80)
    assert a.sync() == b.sync() == 0
    _same(a.state(), b.state(), "compact vs direct")


def test_lazy_capped_patch_runs_the_capped_pass_in_k_post():
    """A patch compressed to 0.7 dx with K = 29: its queries have more than K in range, so
    every step runs the capped pass inside k_post (then, with a re-sort, a grid barrier
    and the sort).  Matches the oracle, which applies the K-nearest rule."""
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(width=0.5, height=0.6, dx=0.02)
    sc.K = 29
    nf = sc.n_fluid
    x = sc.pos[:nf]
    patch = (x[:, 0] < 0.16) & (x[:, 1] < 0.16)
    sc.pos[:nf][patch] = x[patch] * 0.7
    g = GpuSolver(sc, skin=0.02)
    g.step(40)
    assert g.sync() == 40
    nres = g.resorts()
    o = Sim(sc, nthreads=0)
    o.step(40)
    drho = _rel(g.state(), o.state(), nf)
    _say(f"capped patch: 40 capped steps, {nres} re-sorts, vs oracle rho {drho:.2e}")
    assert drho < 1e-6


def test_lazy_ref_mode_matches_oracle():
    """The reference's run-phase scheme (DynamicPressure transport, fluctuations, viscous
    forces) with re-sorts: 300 steps of the 5K dam break against the oracle."""
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(jitter=0.05, seed=5)
    sc.mode = scenes.MODE_REF
    sc.kappa, sc.mu, sc.eps = 0.015, 1e-3, 1.0
    sc.alpha_av = 0.0
    g = GpuSolver(sc, skin=0.02)
    g.step(300)
    assert g.sync() == 0
    nres = g.resorts()
    o = Sim(sc, nthreads=0)
    o.step(300)
    n = sc.n_fluid
    so, sg = o.state(), g.state()
    drho = _rel(sg, so, n)
    Po, Pg = o.ref()["P"][:n], g.pressure()[:n]
    dP = float(np.max(np.abs(Pg - Po)) / np.max(np.abs(Po)))
    _say(f"ref mode: {nres} re-sorts in 300 steps; vs oracle rho {drho:.2e}, P {dP:.2e} of max")
    assert nres > 1 and drho < 1e-6 and dP < 1e-5


def test_lazy_state_set_from_outside_restarts_the_sort():
    """State read and written back through the device state exchange: the next step
    sorts from scratch, and the run continues as a sort-every-step handle's does (to
    rounding)."""
    from uph_amd.devbuf import DevBuf
    from uph_amd.solver import GpuSolver
    sc = _scene(seed=8)
    n, nf = sc.n, sc.n_fluid
    a, b = GpuSolver(sc, skin=0.02), GpuSolver(sc, sort_every_step=True)
    for g in (a, b):
        g.step(40)
        pos, vel, rho = DevBuf(n * 16), DevBuf(n * 16), DevBuf(n * 8)
        g.state_to_device(pos.ptr, vel.ptr, rho.ptr)
        r = rho.download((n,), np.float64)
        r[:nf] *= 1.0 + 1e-4          # a change the next steps must see
        rho.upload(r)
        g.state_from_device(pos.ptr, vel.ptr, rho.ptr, restart=False)
        g.step(40)
    assert a.sync() == b.sync() == 0
    drho, dxe = _rel(a.state(), b.state(), nf), _dx_err(a.state(), b.state(), nf, sc.dx)
    _say(f"state set after 40 steps: lazy vs sorting every step rho {drho:.2e}, x {dxe:.2e} dx")
    assert drho < 1e-8 and dxe < 1e-6


def test_lazy_default_depends_on_the_batch_size():
    """By default a handle of at most 16 scenes re-sorts lazily and a larger batch every
    step (each of its steps would likely re-sort); sort_every_step=False asks for the
    lazy re-sort whatever the size."""
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(width=0.2, height=0.2, dx=0.02)
    assert GpuSolver([sc] * 16).resorts() >= 0
    big = GpuSolver([sc] * 17)
    assert big.resorts() == -1 and big.skin() == 0.0
    forced = GpuSolver([sc] * 17, sort_every_step=False)
    assert forced.skin() == 0.05
    forced.step(5)
    assert forced.sync() == 0 and forced.resorts() >= 1


def test_lazy_vorticity_and_rates_read_the_last_evaluation():
    """approxVorticity and the rates after a call come from the last step's own
    evaluation: equal, to rounding, to those of a handle that sorts every step."""
    from uph_amd.solver import GpuSolver
    sc = _scene(seed=10)
    a, b = GpuSolver(sc, skin=0.02, keep_rates=True), GpuSolver(sc, sort_every_step=True, keep_rates=True)
    for g in (a, b):
        g.step(57)
    va, vb = a.vorticity(), b.vorticity()
    ra, rb = a.rates(), b.rates()
    scale = float(np.max(np.abs(vb)))
    dv = float(np.max(np.abs(va - vb))) / scale
    dR = float(np.max(np.abs(ra["R"] - rb["R"]))) / float(np.max(np.abs(rb["R"])))
    _say(f"after 57 steps ({a.resorts()} re-sorts): vorticity {dv:.2e}, R {dR:.2e} of max")
    assert dv < 1e-6 and dR < 1e-5
