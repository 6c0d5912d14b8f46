"""The oracle's moving-boundary kinematics (inputbp.c:192-283): the lid, the gate and
the sloshing tank.

Pinned by what the motions must satisfy rather than by their formulas: a wall's
displacement between two calls divided by the time between them is its velocity at
the midpoint (exact for the lid, second order for the sloshing walls, which also
return to the deck at a full period and reach the stroke at half of it); the gate
rises 1.5 (LocalDepth - 15 dx) in all and then rests; the sloshing tank keeps its
width (FlatBottomLength) and its bottom's spacing (dx / 2); rows outside the moving
set never move.  The reference's sloshing loop zeroes the left wall's motion once it
meets a moving row of another itype (inputbp.c:270-275); that is checked on a tank
with such a row in the middle of its moving set.
"""
import math

import numpy as np
import pytest

from uph_amd import scenes


def _walls_at(sc, t_steps):
    """Boundary positions and velocities after each listed number of steps."""
    from oracle.oracle import Sim
    o = Sim(sc, nthreads=4)
    out = {}
    done = 0
    for k in sorted(t_steps):
        o.step(k - done)
        done = k
        st = o.state()
        out[k] = (st["x"][sc.n_fluid:].copy(), st["v"][sc.n_fluid:].copy(), st)
    return out


def _moving(sc):
    return slice(sc.n_fixed, sc.n_fixed + sc.n_moving)


def test_lid_moves_at_lid_velocity():
    sc = scenes.lid_cavity(size=0.1, dx=0.005, lid_u=2.0)
    w = _walls_at(sc, [10, 30])
    mv, fx = _moving(sc), slice(0, sc.n_fixed)
    (x10, v10, _), (x30, v30, st) = w[10], w[30]
    dt = sc.dt
    assert np.allclose((x30[mv, 0] - x10[mv, 0]) / (20 * dt), 2.0, rtol=1e-12)
    assert np.all(v10[mv, 0] == 2.0) and np.all(v30[mv, 0] == 2.0)
    # the lid started from its deck position at t = 0
    assert np.allclose(x30[mv, 0] - sc.pos[sc.n_fluid:][mv, 0], 2.0 * 30 * dt, rtol=1e-12)
    assert np.array_equal(x30[mv, 1], sc.pos[sc.n_fluid:][mv, 1])
    assert np.all(v30[mv, 1] == 0.0)
    assert np.array_equal(x30[fx], sc.pos[sc.n_fluid:][fx])
    # the localVelocity of the lid is its velocity (inputbp.c:221)
    assert np.all(st["vloc"][sc.n_fluid:][mv, 0] == 2.0)


def test_lid_drags_the_top_layer():
    """The lid's shear reaches the fluid: the top rows move along it, the bottom rows
    hardly at all."""
    sc = scenes.lid_cavity(size=0.1, dx=0.005, lid_u=1.0)
    st = _walls_at(sc, [400])[400][2]
    nf = sc.n_fluid
    y = sc.pos[:nf, 1]
    top = y > 0.1 - 2.5 * sc.dx
    bottom = y < 2.5 * sc.dx
    u_top = float(st["v"][:nf][top, 0].mean())
    assert u_top > 0.05
    assert u_top > 5 * abs(float(st["v"][:nf][bottom, 0].mean()))


def _gate(stop_steps=20):
    sc = scenes.gate_dam_break(width=0.2, height=0.3, dx=0.01, tank=(0.6, 0.5), gate_u=2.0)
    # LocalDepth that stops the gate after stop_steps + 1/2 steps
    sc.depth = 15 * sc.dx + 2.0 * (stop_steps + 0.5) * sc.dt / 1.5
    return sc


def test_gate_rises_then_rests():
    sc = _gate(20)
    w = _walls_at(sc, [5, 15, 25, 40])
    mv = _moving(sc)
    y0 = sc.gate_y[sc.n_fluid:][mv]
    rise = 1.5 * (sc.depth - 15 * sc.dx)
    for k in (5, 15):
        x, v, _ = w[k]
        assert np.allclose(x[mv, 1] - y0, 2.0 * k * sc.dt, rtol=1e-12, atol=0)
        assert np.all(v[mv, 1] == 2.0)
    for k in (25, 40):
        x, v, _ = w[k]
        assert np.allclose(x[mv, 1] - y0, rise, rtol=1e-12, atol=0)
        assert np.all(v[mv, 1] == 0.0)
    assert np.allclose(rise, 2.0 * 20.5 * sc.dt, rtol=1e-12)
    # x is the deck's, the fixed tank does not move
    x, _, _ = w[40]
    assert np.array_equal(x[mv, 0], sc.pos[sc.n_fluid:][mv, 0])
    assert np.array_equal(x[:sc.n_fixed], sc.pos[sc.n_fluid:sc.n_fluid + sc.n_fixed])


def test_gate_height_column_is_what_the_gate_follows():
    """tempPositionGate, not the deck y, is the base of the gate's height."""
    sc = _gate(20)
    sc.gate_y = sc.gate_y.copy()
    sc.gate_y[sc.n_fluid + sc.n_fixed:] += 0.25
    x, _, _ = _walls_at(sc, [5])[5]
    mv = _moving(sc)
    assert np.allclose(x[mv, 1] - sc.pos[sc.n_fluid:][mv, 1], 0.25 + 2.0 * 5 * sc.dt, rtol=1e-12)


def _rows(sc):
    it = sc.itype[sc.n_fluid + sc.n_fixed:]
    return it == 2, it == 3, it == 1


def test_sloshing_walls_move_together():
    period = 0.02
    sc = scenes.sloshing_tank(width=0.1, depth=0.05, dx=0.005, wall_height=0.1, stroke=0.01, period=period)
    dt = sc.dt
    ks = [10, 11, 12]
    w = _walls_at(sc, ks)
    mv = _moving(sc)
    L, R, B = _rows(sc)
    for k in ks:
        x, v, _ = w[k]
        xm, vm = x[mv, 0], v[mv, 0]
        # rigid tank: width FlatBottomLength, bottom spacing dx / 2, one velocity
        assert np.allclose(xm[R] - xm[L].mean(), sc.flat_bottom, rtol=0, atol=1e-15)
        assert np.allclose(np.diff(xm[B]), 0.5 * sc.dx, rtol=0, atol=1e-15)
        assert np.allclose(xm[B][0], xm[L][0], rtol=0, atol=1e-15)
        assert np.ptp(vm) == 0.0
        assert np.array_equal(x[mv, 1], sc.pos[sc.n_fluid:][mv, 1])
    # displacement over two steps / 2 dt = velocity at the middle step, to O(dt^2)
    vc = w[11][1][mv, 0][0]
    fd = (w[12][0][mv, 0][0] - w[10][0][mv, 0][0]) / (2 * dt)
    om = 2 * math.pi / period
    assert abs(fd - vc) < (om * dt) ** 2 * 0.5 * sc.m_amp * om
    assert abs(vc) > 0.01 * 0.5 * sc.m_amp * om


def test_sloshing_cycle_extremes():
    """At half a period the walls are one stroke from the deck, at a full period back."""
    sc = scenes.sloshing_tank(width=0.1, depth=0.05, dx=0.005, wall_height=0.1, stroke=0.01)
    n_half = 40
    sc.m_omega = math.pi          # half period = 1 s
    sc.dt = 1.0 / n_half
    w = _walls_at(sc, [n_half, 2 * n_half])
    mv = _moving(sc)
    L, R, B = _rows(sc)
    xh, vh, _ = w[n_half]
    xf, vf, _ = w[2 * n_half]
    assert np.allclose(xh[mv, 0][L], 0.01, atol=1e-15)
    assert np.allclose(xh[mv, 0][R], sc.flat_bottom + 0.01, atol=1e-15)
    assert np.allclose(vh[mv, 0], 0.0, atol=1e-15)
    assert np.allclose(xf[mv, 0][L], 0.0, atol=1e-15)
    assert np.allclose(xf[mv, 0][B], np.arange(B.sum()) * 0.5 * sc.dx, atol=1e-15)


def _halted_tank():
    """A sloshing tank whose moving set holds rows of itype 4 between its bottom and
    its left wall."""
    sc = scenes.sloshing_tank(width=0.1, depth=0.05, dx=0.005, wall_height=0.1, stroke=0.01, period=0.02)
    it = sc.itype[sc.n_fluid + sc.n_fixed:]
    nb = int((it == 1).sum())
    extra = 3
    first = sc.n_fluid + sc.n_fixed + nb
    sc.itype = sc.itype.copy()
    sc.itype[first:first + extra] = 4
    return sc, first - sc.n_fluid, extra


def test_sloshing_loop_zeroes_motion_after_another_itype():
    sc, first, extra = _halted_tank()
    x, v, _ = _walls_at(sc, [11])[11]
    deck = sc.pos[sc.n_fluid:]
    ity = sc.itype[sc.n_fluid:]
    L = 0.5 * sc.m_amp * (1 - math.cos(sc.m_omega * 11 * sc.dt))
    assert L > 1e-4
    bottom = np.arange(sc.n_fixed, first)
    assert np.all(ity[bottom] == 1)
    assert np.allclose(x[bottom, 0], (bottom - sc.n_fixed) * 0.5 * sc.dx + L, atol=1e-15)
    assert np.all(v[bottom, 0] != 0.0)
    other = np.arange(first, first + extra)
    assert np.array_equal(x[other], deck[other]) and np.array_equal(v[other], sc.vel[sc.n_fluid:][other])
    rest = np.arange(first + extra, sc.n - sc.n_fluid)
    left, right = rest[ity[rest] == 2], rest[ity[rest] == 3]
    # left wall: L and its velocity were zeroed; right wall: its position was taken
    # before the loop and keeps L, its velocity is the zeroed one
    assert np.all(x[left, 0] == 0.0) and np.all(v[left, 0] == 0.0)
    assert np.allclose(x[right, 0], sc.flat_bottom + L, atol=1e-15) and np.all(v[right, 0] == 0.0)
    assert len(left) > 0 and len(right) > 0


@pytest.mark.parametrize("make", [scenes.lid_cavity, scenes.gate_dam_break, scenes.sloshing_tank])
def test_deck_directory_round_trip(tmp_path, make):
    """write_deck_dir then load_deck_dir gives back the scene's rows, motion and the
    constants the motion uses."""
    sc = make()
    scenes.write_deck_dir(str(tmp_path), sc)
    back = scenes.load_deck_dir(str(tmp_path))
    assert (back.n_fluid, back.n_fixed, back.n_moving) == (sc.n_fluid, sc.n_fixed, sc.n_moving)
    assert back.motion == sc.motion
    assert np.array_equal(back.itype, sc.itype)
    assert np.allclose(back.pos, sc.pos, atol=1e-10, rtol=0)
    assert np.allclose(back.gate_y, sc.gate_y if sc.gate_y is not None else 0.0, atol=1e-10, rtol=0)
    for k in ("lid_u", "depth", "flat_bottom", "m_amp", "dt", "rho0", "gamma", "g", "gravity"):
        assert math.isclose(getattr(back, k), getattr(sc, k), rel_tol=1e-12, abs_tol=1e-300), k
    assert math.isclose(back.c0, sc.c0, rel_tol=1e-12)
    assert math.isclose(back.h, sc.h, rel_tol=1e-12)
    if sc.motion == scenes.MOTION_SLOSH:
        assert math.isclose(back.m_omega, sc.m_omega, rel_tol=1e-12)
