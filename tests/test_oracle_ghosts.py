"""The oracle's mirror ("virtual") particles, derivatives.c:3193-4192, and their one
effect on the step: the neighbour search ranks them among the K nearest
(derivatives.c:1069-1085) and the pair filter then drops them (Search.cc:33).

Pinned by facts independent of the generator's formulas: on a lattice the number of
mirror rows of each section is a count of lattice columns or rows within reach of its
line; every mirror map is an isometry (distances among a section's rows equal those
among their sources), keeps a row's distance to its line and flips the velocity
component normal to it; the slope's images have their midpoints on the line
y = (x - FlatBottomLength)/15 and their displacement normal to it; rows come section
by section, sources in index order.  The search over particles and mirror rows agrees
with the brute-force statement of the SearchNeighbors contract, and where K binds the
mirror rows displace real neighbours.
"""
import math

import numpy as np
import pytest

from uph_amd import scenes


def _tank(nx, ny, dx=0.02, vtype=8, h_factor=1.3, jitter=0.0, seed=0, fill_width=True):
    """Fluid lattice [0, nx dx] x [0, ny dx] in a tank as wide as the fluid (FlatBottomLength)."""
    W = nx * dx
    sc = scenes.dam_break(width=W, height=ny * dx, dx=dx, tank=(W if fill_width else 2 * W, 2 * ny * dx),
                          h_factor=h_factor, jitter=jitter, seed=seed)
    sc.flat_bottom = W if fill_width else 2 * W
    sc.depth = ny * dx
    sc.vtype = vtype
    return sc


def _ghosts_after_one_call(sc, rng_v=None):
    from oracle.oracle import Sim
    if rng_v is not None:
        nf = sc.n_fluid
        sc.vel[:nf] = rng_v.normal(0, 0.5, (nf, 2))
        sc.lvel[:nf] = sc.vel[:nf] + rng_v.normal(0, 0.05, (nf, 2))
    o = Sim(sc, nthreads=2)
    o.derivatives(sc.dt)
    return o, o.ghosts()


def _sections(vtype):
    from oracle.oracle import lib
    sec = np.zeros(8, np.int32)
    m = lib().orc_ghost_sections(vtype, sec.ctypes.data)
    return list(sec[:m])


def _split(sc, g, vtype):
    """Ghost rows of each section, with their sources (found by the section's own
    geometry: the mirror of a ghost is a fluid position)."""
    nf = sc.n_fluid
    src_pos = sc.pos[:nf]
    out = []
    k = 0
    lines = _sections(vtype)
    for sec in lines:
        rows = []
        while k < len(g["x"]):
            gx, gy = g["x"][k]
            cand = {1: (-gx, gy), 2: (-gx, gy), 3: (gx, -gy), 4: (-gx, -gy),
                    6: (2 * sc.flat_bottom - gx, gy), 9: (gx, None), 8: (gx, None)}.get(sec)
            if cand is None:
                break
            if cand[1] is None:
                # top sections: the source is the fluid row whose image this is
                d = (gy - sc.depth) / 2.0
                y_src = sc.depth + d + (sc.dx if sec == 8 else 0.0)
                y_alt = sc.depth - d + (sc.dx if sec == 8 else 0.0)
                hit = np.nonzero((np.abs(src_pos[:, 0] - gx) < 1e-12) &
                                 ((np.abs(src_pos[:, 1] - y_src) < 1e-9) | (np.abs(src_pos[:, 1] - y_alt) < 1e-9)))[0]
            else:
                hit = np.nonzero(np.all(np.abs(src_pos - np.array(cand)) < 1e-12, axis=1))[0]
            if len(hit) == 0 or (rows and hit[0] <= rows[-1][1]):
                break
            rows.append((k, int(hit[0])))
            k += 1
        out.append((sec, rows))
    return out, k


def test_section_lists_of_each_virtual_type():
    assert _sections(1) == [2, 3, 4, 10]
    assert _sections(3) == [1, 3, 8, 4, 6]
    assert _sections(4) == _sections(8) == [1, 3, 4, 6]
    assert _sections(5) == [2, 4, 5, 7]
    assert _sections(7) == [1, 3, 9, 4, 6]
    for vt in (0, 2, 6, 77):
        assert _sections(vt) == []


def test_no_mirror_rows_without_a_virtual_type():
    for vt in (0, 2, 77):
        _, g = _ghosts_after_one_call(_tank(10, 8, vtype=vt))
        assert len(g["x"]) == 0


def test_lattice_counts_of_the_dam_break_sections():
    """vtype 8 on a 10 x 8 lattice, h = 1.3 dx: columns / rows with centres closer than
    2h = 2.6 dx to a line are 3 (0.5, 1.5, 2.5 dx)."""
    nx, ny = 10, 8
    sc = _tank(nx, ny, vtype=8)
    _, g = _ghosts_after_one_call(sc)
    parts, used = _split(sc, g, 8)
    assert used == len(g["x"])
    counts = [len(rows) for _, rows in parts]
    assert counts == [3 * ny, 3 * nx, 3 * 3, 3 * ny]
    # every section lists its sources in index order
    for _, rows in parts:
        src = [s for _, s in rows]
        assert src == sorted(src)


def test_hydrostatic_tank_top_section():
    """vtype 7's top images sit at LocalDepth + 2 |y - LocalDepth|: for the top three
    lattice rows (0.5, 1.5, 2.5 dx below the surface) at 1, 3, 5 dx above it."""
    nx, ny = 6, 9
    sc = _tank(nx, ny, vtype=7)
    _, g = _ghosts_after_one_call(sc)
    parts, used = _split(sc, g, 7)
    assert used == len(g["x"])
    sec, rows = parts[2]
    assert sec == 9 and len(rows) == 3 * nx
    heights = sorted({round((g["x"][k, 1] - sc.depth) / sc.dx, 9) for k, _ in rows})
    assert heights == [1.0, 3.0, 5.0]


def test_mirror_maps_are_isometries_that_flip_the_normal_velocity():
    sc = _tank(12, 9, vtype=8, jitter=0.15, seed=3)
    _, g = _ghosts_after_one_call(sc, np.random.default_rng(5))
    nf = sc.n_fluid
    # sources recovered with a tolerance: jittered positions map exactly, compare
    # against the fluid rows of the same call
    from oracle.oracle import Sim
    o = Sim(sc, nthreads=1)
    o.derivatives(sc.dt)
    st = o.state()
    x, v, vl = st["x"][:nf], st["v"][:nf], st["vloc"][:nf]
    k = 0
    for sec, line in [(1, 0.0), (3, 0.0), (4, 0.0), (6, sc.flat_bottom)]:
        if sec == 1:
            src = np.nonzero(np.abs(x[:, 0]) < 2 * sc.h)[0]
            n_hat = np.array([1.0, 0.0])
        elif sec == 3:
            src = np.nonzero((x[:, 1] < 2 * sc.h) & (x[:, 0] <= sc.flat_bottom))[0]
            n_hat = np.array([0.0, 1.0])
        elif sec == 4:
            src = np.nonzero((np.abs(x[:, 0]) < 2 * sc.h) & (x[:, 1] < 2 * sc.h))[0]
            n_hat = None
        else:
            src = np.nonzero(np.abs(x[:, 0] - line) < 2 * sc.h)[0]
            n_hat = np.array([1.0, 0.0])
        gx = g["x"][k:k + len(src)]
        gv, gl = g["v"][k:k + len(src)], g["vloc"][k:k + len(src)]
        k += len(src)
        assert len(gx) == len(src) > 0
        d_src = np.linalg.norm(x[src][:, None] - x[src][None], axis=2)
        d_gh = np.linalg.norm(gx[:, None] - gx[None], axis=2)
        assert np.allclose(d_src, d_gh, atol=1e-14)
        if n_hat is not None:
            # distance to the line kept, normal velocity flipped, tangential kept
            on = 0 if n_hat[0] else 1
            assert np.allclose(np.abs(gx[:, on] - line), np.abs(x[src][:, on] - line), atol=1e-15)
            assert np.allclose(gv @ n_hat, -(v[src] @ n_hat), atol=1e-15)
            t_hat = n_hat[::-1]
            assert np.allclose(gv @ t_hat, v[src] @ t_hat, atol=1e-15)
            assert np.allclose(gl @ n_hat, -(vl[src] @ n_hat), atol=1e-15)
        else:
            assert np.allclose(gx, -x[src], atol=1e-15)
            assert np.allclose(gv, -v[src], atol=1e-15)
        assert np.array_equal(g["rho"][k - len(src):k], st["rho"][src])
        assert np.array_equal(g["mass"][k - len(src):k], sc.mass[src])
    assert k == len(g["x"])


def test_slope_images_reflect_in_the_beach_line():
    """virtualType 1: rows beyond FlatBottomLength within 3h of y = (x - FBL)/15 are
    reflected in that line."""
    dx = 0.02
    sc = _tank(30, 6, dx=dx, vtype=1, fill_width=False)
    sc.flat_bottom = 0.3
    # lift the fluid beyond x = 0.3 onto the slope so that some rows are within 3h
    nf = sc.n_fluid
    xs = sc.pos[:nf, 0]
    beyond = xs > 0.3
    sc.pos[:nf, 1][beyond] += (xs[beyond] - 0.3) / 15.0
    o, g = _ghosts_after_one_call(sc, np.random.default_rng(2))
    st = o.state()
    x, v = st["x"][:nf], st["v"][:nf]
    m = 1.0 / 15.0
    dist = np.abs(m * x[:, 0] - x[:, 1] - m * 0.3) / math.sqrt(1 + m * m)
    src = np.nonzero((dist < 3 * sc.h) & (x[:, 0] > 0.3))[0]
    assert len(src) > 10
    gx, gv = g["x"][-len(src):], g["v"][-len(src):]
    mid = 0.5 * (gx + x[src])
    assert np.allclose(mid[:, 1], (mid[:, 0] - 0.3) * m, atol=1e-13)
    disp = gx - x[src]
    assert np.allclose(disp @ np.array([1.0, m]), 0.0, atol=1e-13)
    n_hat = np.array([m, -1.0]) / math.sqrt(1 + m * m)
    assert np.allclose(gv @ n_hat, -(v[src] @ n_hat), atol=1e-13)
    t_hat = np.array([1.0, m]) / math.sqrt(1 + m * m)
    assert np.allclose(gv @ t_hat, v[src] @ t_hat, atol=1e-13)


def test_sloshing_images_follow_the_walls():
    """virtualType 5: the side images are taken about LeftBoundaryPosition = L(t) and
    RightBoundaryPosition = FlatBottomLength + L(t) of the same call."""
    sc = scenes.sloshing_tank(width=0.1, depth=0.05, dx=0.005, wall_height=0.1, stroke=0.01, period=0.02)
    from oracle.oracle import Sim
    o = Sim(sc, nthreads=2)
    t = 7 * sc.dt
    o.derivatives(t)          # images of the positions this call sees
    g = o.ghosts()
    st = o.state()
    L = 0.5 * sc.m_amp * (1 - math.cos(sc.m_omega * t))
    nf = sc.n_fluid
    x = st["x"][:nf]
    nl = int((np.abs(x[:, 0] - L) < 2 * sc.h).sum())
    assert nl > 0
    assert np.allclose(g["x"][:nl, 0], L - np.abs(x[np.abs(x[:, 0] - L) < 2 * sc.h][:, 0] - L), atol=1e-15)
    R = sc.flat_bottom + L
    nr = int((np.abs(x[:, 0] - R) < 2 * sc.h).sum())
    assert nr > 0
    assert np.allclose(g["x"][-nr:, 0], R + np.abs(x[np.abs(x[:, 0] - R) < 2 * sc.h][:, 0] - R), atol=1e-15)


def _pairs(pos, nf, n, pb, K, cutoff, brute):
    from oracle.oracle import lib
    cap = nf * K + 16
    ip = np.zeros(cap, np.int32); jp = np.zeros(cap, np.int32); r2 = np.zeros(cap)
    p = np.ascontiguousarray(pos, np.float64).reshape(-1)
    if brute:
        m = lib().orc_search_brute(p, ip, jp, r2, nf, n, pb, K, cutoff, cap)
    else:
        m = lib().orc_search(p, ip, jp, r2, nf, n, pb, K, cutoff, cap, 2)
    assert m >= 0
    return set(zip(ip[:m].tolist(), jp[:m].tolist()))


def test_search_over_mirror_rows_matches_brute_force_and_they_displace_neighbours():
    """h = 2.691 dx (the reference dambreak deck's nearNeighbor 91): about 88 points in
    range, so K = 40 binds and mirror rows, ranked but never paired, change the sets."""
    sc = _tank(14, 10, vtype=8, h_factor=2.691, jitter=0.05, seed=9)
    o, g = _ghosts_after_one_call(sc)
    nf, n = sc.n_fluid, sc.n
    ng = len(g["x"])
    assert ng > 0
    pos = np.concatenate([o.state()["x"], g["x"]])
    pb = nf + sc.n_fixed + sc.n_moving - 1
    cutoff = sc.h * sc.h * 4
    with_g = _pairs(pos, nf, n + ng, pb, 40, cutoff, brute=False)
    assert with_g == _pairs(pos, nf, n + ng, pb, 40, cutoff, brute=True)
    assert o.niac == len(with_g)
    without = _pairs(pos[:n], nf, n, pb, 40, cutoff, brute=True)
    assert with_g != without
    assert with_g < without or len(with_g) < len(without)
    # only queries within reach of a mirror line lose neighbours
    changed = {i for i, _ in without - with_g}
    x = pos[:nf]
    reach = 4 * sc.h
    near = (x[:, 0] < reach) | (x[:, 1] < reach) | (np.abs(x[:, 0] - sc.flat_bottom) < reach)
    assert changed and all(near[i] for i in changed)
