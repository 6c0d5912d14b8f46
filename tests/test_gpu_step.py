"""The GPU stepping path against the FP64 oracle.

Pair sums run in FP32 with compensated accumulation on the GPU and in FP64 in
the oracle; positions, velocities and densities are integrated in FP64 on
both.  Tolerances (stated per test) follow the precision plan: rates to 1e-5
of their largest magnitude, density to 1e-5 relative after 1000 steps of the
5,000-particle dam break.
"""
import numpy as np
import pytest

from conftest import load_golden_scene
from uph_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-step] {msg}", flush=True)


def _rel_rho(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def _moving_scene(sc, seed=1, amp=0.3):
    rng = np.random.default_rng(seed)
    nf = sc.n_fluid
    sc.vel[:nf] = rng.normal(0, amp, (nf, 2))
    sc.lvel[:nf] = sc.vel[:nf] + rng.normal(0, 0.1 * amp, (nf, 2))
    return sc


def _compare_rates(sc, name):
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    _say(f"rates {name}: n={sc.n}")
    g = GpuSolver(sc, keep_rates=True)
    g.step(1)
    g.sync()
    rg = g.rates()
    o = Sim(sc, nthreads=0)
    o.derivatives(sc.dt)
    ro = o.rates()
    nf = sc.n_fluid
    out = {}
    for key, okey in (("R", "R"), ("Rt", "Rt"), ("a", "a"), ("at", "at")):
        ref = ro[okey][:nf]
        err = np.abs(rg[key].astype(np.float64) - ref).max() / max(np.abs(ref).max(), 1e-300)
        out[okey] = err
    _say(f"rates {name}: rel-to-max errors {out}")
    return out, g


def test_rates_jittered_moving_fluid():
    sc = _moving_scene(scenes.dam_break(jitter=0.1, seed=4))
    err, _ = _compare_rates(sc, "dam break 5K jittered, random v")
    for k, e in err.items():
        assert e < 1e-5, (k, e)


def test_rates_reference_hydrostatic_deck():
    sc = load_golden_scene("hydrostatic_init.npz", alpha_av=0.1)
    err, _ = _compare_rates(sc, "tests/hydrostaticPressure FluidParticleInit")
    for k, e in err.items():
        assert e < 1e-5, (k, e)


def test_rates_capped_neighbour_sets():
    """tests/dambreak deck: ~88 particles in range, so the K = 40 path is taken."""
    sc = _moving_scene(load_golden_scene("dambreak.npz"), amp=0.1)
    err, g = _compare_rates(sc, "tests/dambreak (K=40 binds)")
    assert g.sync() == 1                      # the capped path ran in the one step
    for k, e in err.items():
        assert e < 1e-5, (k, e)


def test_rates_capped_rows_longer_than_128_slots():
    """h = 4 dx: about 200 particles in range and candidate rows of up to ~190 slots, so the
    capped pass takes its direct loop for rows past 128 slots and the bit walk for the
    rest (near the walls); both against the oracle, and the compacted kernel against the
    direct one bit for bit over 10 steps."""
    from uph_amd.solver import GpuSolver
    sc = _moving_scene(scenes.dam_break(width=0.4, height=0.4, dx=0.02, tank=(0.8, 0.8), h_factor=4.0,
                                        jitter=0.1, seed=21), amp=0.1)
    err, g = _compare_rates(sc, "h = 4 dx, rows past 128 slots (K=40 binds)")
    assert g.sync() == 1
    for k, e in err.items():
        assert e < 1e-5, (k, e)
    a = GpuSolver(sc, pair_variant=0)
    b = GpuSolver(sc, pair_variant=1)
    a.step(10)
    b.step(10)
    assert a.sync() == 10 and b.sync() == 10
    sa, sb = a.state(), b.state()
    for key in ("x", "v", "rho"):
        assert np.array_equal(sa[key], sb[key]), key


def _run_both(sc, steps, name, nthreads=0, **gopts):
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    _say(f"{name}: n={sc.n} steps={steps}")
    g = GpuSolver(sc, **gopts)
    g.step(steps)
    g.sync()
    sg = g.state()
    o = Sim(sc, nthreads=nthreads)
    o.step(steps)
    so = o.state()
    nf = sc.n_fluid
    drho = _rel_rho(sg["rho"][:nf], so["rho"][:nf])
    dx = float(np.abs(sg["x"][:nf] - so["x"][:nf]).max()) / sc.dx
    dv = float(np.abs(sg["v"][:nf] - so["v"][:nf]).max())
    _say(f"{name}: max rel rho err {drho:.3e}, max |dx|/dx {dx:.3e}, max |dv| {dv:.3e}")
    return drho, dx, sg, so, g


def test_dam_break_5k_1000_steps():
    """Dam break 1.0 x 2.0 m, 5,000 fluid, dx = 0.02, 1000 steps."""
    sc = scenes.dam_break()
    drho, dx, sg, so, g = _run_both(sc, 1000, "dam break 5K 1000 steps")
    assert g.sync() == 0
    assert drho < 1e-5
    assert dx < 1e-3


def test_reference_hydrostatic_deck_200_steps():
    sc = load_golden_scene("hydrostatic_init.npz", alpha_av=0.1)
    drho, dx, *_ = _run_both(sc, 200, "tests/hydrostaticPressure 200 steps")
    assert drho < 1e-5
    assert dx < 1e-3


def test_piston_wave_tank():
    """Moving boundary (sinusoidal piston) and a sloping beach."""
    sc = scenes.wave_tank(length=1.0, depth=0.3, dx=0.01, beach_toe=0.7, amp=0.02, period=0.5)
    drho, dx, sg, so, _ = _run_both(sc, 300, "wave tank piston 300 steps")
    assert drho < 1e-5
    assert dx < 1e-3
    # the piston particles are where the 300th derivatives() call put them
    mv = slice(sc.n_fluid + sc.n_fixed, sc.n)
    assert np.allclose(sg["x"][mv, 1], sc.pos[mv, 1])
    assert float(np.abs(sg["x"][mv] - so["x"][mv]).max()) < 1e-12
    assert float(np.abs(sg["v"][mv] - so["v"][mv]).max()) < 1e-12


def test_reference_waves_deck_piston():
    """The reference's waves deck with its own piston (inputbp.c:194-207)."""
    sc = load_golden_scene("waves.npz", alpha_av=0.1)
    drho, dx, *_ = _run_both(sc, 100, "tests/waves reference piston 100 steps")
    assert drho < 1e-5
    assert dx < 1e-3


def test_batch_equals_single_scenes():
    from uph_amd.solver import GpuSolver
    scs = [scenes.dam_break(width=0.4, height=0.5, dx=0.02),
           scenes.dam_break(width=0.3, height=0.6, dx=0.015),
           scenes.dam_break(width=0.5, height=0.3, dx=0.025, jitter=0.05, seed=1)]
    _say("batch of 3 vs single")
    gb = GpuSolver(scs)
    gb.step(50)
    for k, sc in enumerate(scs):
        gs = GpuSolver(sc, skin=gb.skin())   # a batch's default cell growth differs from one scene's
        gs.step(50)
        a, b = gb.state(k), gs.state(0)
        for key in ("x", "v", "rho"):
            assert np.array_equal(a[key], b[key]), (k, key)


def test_graph_replay_equals_eager():
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(width=0.5, height=0.5, dx=0.02, jitter=0.05, seed=9)
    _say("graph vs eager 40 steps")
    ge = GpuSolver(sc, sort_every_step=True)   # a captured step sorts every step
    gg = GpuSolver(sc, use_graph=True)
    for chunk in (1, 9, 30):
        ge.step(chunk)
        gg.step(chunk)
        a, b = ge.state(), gg.state()     # eager work and syncs between replays
        for key in ("x", "v", "rho"):
            assert np.array_equal(a[key], b[key]), (chunk, key)


def test_free_fall_matches_closed_form():
    from uph_amd.solver import GpuSolver
    N = 3
    pos = np.array([[0.0, 0.0], [10.0, 0.0], [20.0, 5.0]])
    vel = np.array([[0.5, 1.0], [0.0, 0.0], [-1.0, 2.0]])
    sc = scenes.Scene(pos=pos, vel=vel, rho=np.full(N, 1000.0), pressure=np.zeros(N),
                      mass=np.full(N, 0.1), itype=np.ones(N, np.int32), lvel=vel.copy(),
                      n_fluid=N, n_fixed=0, n_moving=0, dx=0.01, h=0.013, c0=40.0,
                      B=1000 * 1600 / 7, gravity=1, alpha_av=0.1, dt=1e-3,
                      box=(-1, -50, 30, 10))
    _say("free fall")
    g = GpuSolver(sc)
    n = 500
    g.step(n)
    st = g.state()
    t = n * sc.dt
    xe = pos + vel * t + np.array([0.0, -0.5 * 9.81 * t * t])
    assert np.allclose(st["x"], xe, rtol=0, atol=1e-11)


def test_errors_are_reported():
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(width=0.2, height=0.2)
    sc.rho[sc.n_fluid] = 0.0                # wall densities must be positive
    with pytest.raises(RuntimeError):
        GpuSolver(sc)
    g = GpuSolver(scenes.dam_break(width=0.2, height=0.2))
    with pytest.raises(IndexError):
        g.state(5)
    # the C entry point itself rejects the index
    import ctypes
    from uph_amd import solver
    x = np.zeros(8)
    assert solver.lib().uph_get_state(g.h, 5, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      None, None) == -1
    assert b"out of range" in solver.lib().uph_last_error(g.h)
    with pytest.raises(RuntimeError):
        g.rates()                           # created without keep_rates


def _tait_p(sc, rho):
    return sc.B * ((rho / sc.rho0) ** 7 - 1.0)


def test_dam_break_1m_200_steps():
    """Dam break 1.0 x 2.0 m at dx = sqrt(2) mm (999,698 fluid + 21,228 wall): the GPU
    (FP32 pair terms, FP64 sums and integration) against the FP64 oracle after 200 steps.
    Tolerances: density 1e-5 relative (max over fluid, matched by original index),
    positions 1e-3 dx; the Tait pressure's error is reported relative to its largest
    magnitude."""
    sc = scenes.dam_break(dx=2 ** 0.5 * 1e-3)
    drho, dx, sg, so, g = _run_both(sc, 200, "dam break 1M, 200 steps")
    nf = sc.n_fluid
    pg, po = _tait_p(sc, sg["rho"][:nf]), _tait_p(sc, so["rho"][:nf])
    dp = float(np.abs(pg - po).max() / np.abs(po).max())
    _say(f"dam break 1M after 200 steps: rho [{sg['rho'][:nf].min():.3f}, {sg['rho'][:nf].max():.3f}], "
         f"P err {dp:.3e} of max")
    assert g.sync() == 0
    assert np.all(np.isfinite(sg["x"])) and np.all(np.isfinite(sg["rho"]))
    assert drho < 1e-5
    assert dx < 1e-3
    assert dp < 1e-4


def test_dam_break_250k_1000_steps():
    """The config-2 dam break at twice its spacing (dx = 2 sqrt(2) mm, 249,571 fluid):
    the north star's density bar, 1e-5 relative at 1000 steps, at a quarter of the 1M
    workload (the largest the FP64 oracle runs in about half a minute)."""
    sc = scenes.dam_break(dx=2 * 2 ** 0.5 * 1e-3)
    drho, dx, sg, so, g = _run_both(sc, 1000, f"dam break {sc.n_fluid} fluid, 1000 steps")
    assert g.sync() == 0
    assert drho < 1e-5
    assert dx < 1e-3


@pytest.mark.parametrize("name", ["jittered", "capped"])
def test_compact_variant_equals_direct(name):
    """The LDS-compacted pair kernel and the direct one give identical results."""
    from uph_amd.solver import GpuSolver
    if name == "jittered":
        sc = _moving_scene(scenes.dam_break(width=0.5, height=0.8, jitter=0.1, seed=12))
    else:
        sc = _moving_scene(load_golden_scene("dambreak.npz"), amp=0.1)
    _say(f"compact vs direct: {name}")
    a = GpuSolver(sc, pair_variant=0)
    b = GpuSolver(sc, pair_variant=1)
    a.step(20)
    b.step(20)
    sa, sb = a.state(), b.state()
    for key in ("x", "v", "rho"):
        assert np.array_equal(sa[key], sb[key]), key


def test_rates_hydrostatic_walls():
    """Walls whose deck density differs from rho0 (localDensity/density != 1)."""
    sc = _moving_scene(scenes.wave_tank(length=1.0, depth=0.3, dx=0.01, beach_toe=0.7, amp=0.02,
                                        period=0.5), amp=0.1)
    err, _ = _compare_rates(sc, "wave tank, hydrostatic walls")
    for k, e in err.items():
        assert e < 1e-5, (k, e)


# ---------------------------------------------------- PolyOrderSeven kernel

def _poly7(sc):
    sc.kernel = scenes.KERNEL_POLY7
    return sc


def test_rates_poly7_jittered_moving_fluid():
    """The PolyOrderSeven pair (derivatives.c:767-823) in the FP32 pair sums."""
    sc = _poly7(_moving_scene(scenes.dam_break(jitter=0.1, seed=4)))
    err, _ = _compare_rates(sc, "PolyOrderSeven, dam break 5K jittered, random v")
    for k, e in err.items():
        assert e < 1e-5, (k, e)


def test_rates_poly7_reference_dambreak_deck():
    """The reference's own kernel on its tests/dambreak deck (h = 2.691 dx, K = 40 binds)."""
    sc = _poly7(_moving_scene(load_golden_scene("dambreak.npz"), amp=0.1))
    err, g = _compare_rates(sc, "PolyOrderSeven, tests/dambreak (K=40 binds)")
    assert g.sync() == 1
    for k, e in err.items():
        assert e < 1e-5, (k, e)


def test_poly7_dam_break_5k_500_steps():
    sc = _poly7(scenes.dam_break())
    drho, dx, sg, so, g = _run_both(sc, 500, "PolyOrderSeven dam break 5K 500 steps")
    assert drho < 1e-5
    assert dx < 1e-3


def test_poly7_compact_variant_equals_direct():
    from uph_amd.solver import GpuSolver
    sc = _poly7(_moving_scene(scenes.dam_break(width=0.5, height=0.8, jitter=0.1, seed=12)))
    _say("PolyOrderSeven compact vs direct")
    a = GpuSolver(sc, pair_variant=0)
    b = GpuSolver(sc, pair_variant=1)
    a.step(20)
    b.step(20)
    sa, sb = a.state(), b.state()
    for key in ("x", "v", "rho"):
        assert np.array_equal(sa[key], sb[key]), key


def test_kernel_must_be_uniform_across_a_batch():
    from uph_amd.solver import GpuSolver
    a = scenes.dam_break(width=0.2, height=0.2)
    b = _poly7(scenes.dam_break(width=0.2, height=0.3))
    with pytest.raises(RuntimeError, match="kernel"):
        GpuSolver([a, b])
    bad = scenes.dam_break(width=0.2, height=0.2)
    bad.kernel = 7
    with pytest.raises(RuntimeError, match="kernel"):
        GpuSolver(bad)


def test_broken_state_is_reported_not_ground_through():
    """70,000 fluid particles stacked inside one cell (as escaped particles of a run that
    has blown up pile into the clamped edge cells): every query has more than 65,536
    candidates, so the step flags the state instead of running an O(N) search per
    particle, and sync() raises."""
    import time
    from uph_amd.solver import GpuSolver
    rng = np.random.default_rng(0)
    n = 70_000
    pos = np.vstack([rng.uniform(0.5, 0.5 + 1e-4, (n, 2)), [[0.0, 0.0], [1.0, 1.0]]])
    N = len(pos)
    c0, B = scenes.tait_constants(1000.0, 7.0, 9.81, 1.0)
    sc = scenes.Scene(pos=pos, vel=np.zeros((N, 2)), rho=np.full(N, 1000.0), pressure=np.zeros(N),
                      mass=np.full(N, 0.01), itype=np.array([1] * n + [2, 2], np.int32),
                      lvel=np.zeros((N, 2)), n_fluid=n, n_fixed=2, n_moving=0, dx=0.01, h=0.013,
                      c0=c0, B=B, dt=1e-5, box=(0.0, 0.0, 1.0, 1.0))
    _say("broken state: 70,000 particles in one cell")
    g = GpuSolver(sc)
    t = time.perf_counter()
    g.step(1)
    with pytest.raises(RuntimeError, match="blown up"):
        g.sync()
    assert time.perf_counter() - t < 10.0
