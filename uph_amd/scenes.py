"""Scene builders and the 12-column particle-deck format.

A scene is the particle set plus the physical constants of one SPH run, laid
out in the reference's array order: fluid particles first, then solid (always
none here), fixed boundary, moving boundary (`System.c:139-195`,
`inputbp.c:71-150`).  The deck format is the reference's: a count line, then
one row per particle of

    x y vx vy density DynamicPressure mass tempPosition tempPositionGate itype localVx localVy

(`input.c:57-72`, `inputbp.c:88-102`).  itype: 1 fluid, 2 fixed wall,
3 moving wall (piston), 5 spring-active wall (`derivatives.c:1415`).

Constants follow the reference's derivations (`System.c:292-313`) where the
scene is one of the reference's own decks; the dam-break and wave-tank builders
use the scheme of the hot path this package implements: cubic-spline kernel
with h = 1.3 dx, Tait equation of state with c0 = 10 sqrt(2 g H),
B = rho0 c0^2 / gamma, Monaghan artificial viscosity, dt = CFL h / c0.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional

import numpy as np

FLUID, FIXED, MOVING, SPRING_WALL = 1, 2, 3, 5

# moving-boundary kinematics
MOTION_NONE = 0
MOTION_PISTON_REF = 1     # x = S/2 (1 - cos wt), v = S w/2 sin wt, absolute (inputbp.c:194-207)
MOTION_SINE_OFFSET = 2    # x = x_deck + A sin wt, v = A w cos wt
MOTION_LID = 3            # x = x_deck + U t (movingBoundary 3, inputbp.c:209-223)
MOTION_GATE = 4           # y = gate_y + U min(t, t_stop) (movingBoundary 2, inputbp.c:225-243)
MOTION_SLOSH = 5          # sloshing tank walls by deck itype (virtualType 5, inputbp.c:244-276)

# smoothing kernel of the pair sums
KERNEL_CUBIC = 0          # cubic spline, Kernel (derivatives.c:44-66)
KERNEL_POLY7 = 1          # PolyOrderSeven pair of CalculateKernel (derivatives.c:767-823, 1099-1101)

# scheme
MODE_CONFIG = 0           # Tait pressure in the momentum terms, artificial viscosity, no fluctuations
MODE_REF = 1              # the reference's run phase: pressure transport, viscous forces, fluctuations


@dataclasses.dataclass
class Scene:
    """Particles (reference order) and constants of one run."""
    pos: np.ndarray          # (N, 2) float64
    vel: np.ndarray          # (N, 2) float64
    rho: np.ndarray          # (N,) float64
    pressure: np.ndarray     # (N,) float64, deck DynamicPressure column
    mass: np.ndarray         # (N,) float64
    itype: np.ndarray        # (N,) int32
    lvel: np.ndarray         # (N, 2) float64, deck localVelocity columns
    n_fluid: int
    n_fixed: int
    n_moving: int
    dx: float
    h: float
    rho0: float = 1000.0
    gamma: float = 7.0
    c0: float = 0.0
    B: float = 0.0
    g: float = 9.81
    gravity: int = 1
    alpha_av: float = 0.1
    k_spring: float = 0.0
    c_damp: float = 0.0
    dt: float = 0.0
    K: int = 40
    motion: int = MOTION_NONE
    m_amp: float = 0.0
    m_omega: float = 0.0
    box: tuple = (0.0, 0.0, 1.0, 1.0)     # (xmin, ymin, xmax, ymax) of the cell grid
    kernel: int = KERNEL_CUBIC
    mode: int = MODE_CONFIG
    kappa: float = 0.0       # thermalDiffusivity0 (ref mode)
    mu: float = 0.0          # viscosity0 (ref mode)
    eps: float = 0.0         # epsilon, weight of the velocity fluctuation (ref mode)
    lid_u: float = 0.0       # lidVelocity (lid and gate)
    depth: float = 0.0       # LocalDepth (the gate's stop time 1.5 (depth - 15 dx) / lid_u)
    vtype: int = 0           # virtualType: mirror particles ranked by the search (derivatives.c:3193-4192)
    stress_type: int = 1     # stressType (ref mode): 2 takes ClassicShearStress + ViscousForcesTwoStep
    flat_bottom: float = 0.0  # FlatBottomLength (the sloshing tank's right wall)
    gate_y: Optional[np.ndarray] = None   # (N,) deck column 9, tempPositionGate
    name: str = "scene"

    @property
    def n(self) -> int:
        return self.n_fluid + self.n_fixed + self.n_moving

    @property
    def n_boundary(self) -> int:
        return self.n_fixed + self.n_moving

    @property
    def cutoff2(self) -> float:
        """Squared search radius, `hsml[0]*hsml[0]*4` (derivatives.c:1086)."""
        return self.h * self.h * 4

    @property
    def particle_boundary(self) -> int:
        """Exclusive bound on neighbour index, Nf+Ns+Nb-1 (derivatives.c:1081-1083)."""
        return self.n_fluid + self.n_boundary - 1

    def copy(self) -> "Scene":
        return dataclasses.replace(
            self, pos=self.pos.copy(), vel=self.vel.copy(), rho=self.rho.copy(),
            pressure=self.pressure.copy(), mass=self.mass.copy(), itype=self.itype.copy(),
            lvel=self.lvel.copy(), gate_y=None if self.gate_y is None else self.gate_y.copy())


def tait_constants(rho0: float, gamma: float, g: float, H: float):
    c0 = 10.0 * math.sqrt(2.0 * g * H)
    B = rho0 * c0 * c0 / gamma
    return c0, B


def hydrostatic_density(y, H, rho0, gamma, g, B):
    """Inverse Tait of p = rho0 g (H - y): rho = rho0 (1 + p/B)^(1/gamma)."""
    p = rho0 * g * (H - y)
    return rho0 * (1.0 + p / B) ** (1.0 / gamma), p


def _wall_params(mass: float, dx: float, g: float):
    """(springConstant0, dampingConstant0) of the generated scenes.

    The reference's spring term k0 (r0 - r) dx/(r m) is subtracted with dx = x_fluid -
    x_wall (derivatives.c:2433-2437), i.e. it pulls fluid toward the wall; walls here
    repel through the pressure term instead, so k0 = 0.  The damping constant keeps the
    reference dambreak deck's ratio to the particle mass (0.02912 / 0.036).
    """
    return 0.0, (0.02912 / 0.036) * mass


def _assemble(fluid_xy, fluid_rho, fluid_p, walls_fixed, walls_moving, mass, **kw) -> Scene:
    nf = len(fluid_xy)
    nfx = len(walls_fixed)
    nmv = len(walls_moving)
    n = nf + nfx + nmv
    pos = np.zeros((n, 2))
    pos[:nf] = fluid_xy
    if nfx:
        pos[nf:nf + nfx] = walls_fixed
    if nmv:
        pos[nf + nfx:] = walls_moving
    rho = np.full(n, kw.get("rho0", 1000.0))
    rho[:nf] = fluid_rho
    pres = np.zeros(n)
    pres[:nf] = fluid_p
    itype = np.empty(n, np.int32)
    itype[:nf] = FLUID
    itype[nf:nf + nfx] = FIXED
    itype[nf + nfx:] = MOVING
    return Scene(pos=pos, vel=np.zeros((n, 2)), rho=rho, pressure=pres,
                 mass=np.full(n, mass), itype=itype, lvel=np.zeros((n, 2)),
                 n_fluid=nf, n_fixed=nfx, n_moving=nmv, **kw)


def _lattice(nx, ny, dx, x0=0.0, y0=0.0):
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    x = x0 + (i.ravel() + 0.5) * dx
    y = y0 + (j.ravel() + 0.5) * dx
    return np.stack([x, y], axis=1)


def _tank_walls(dx, tank_w, tank_h, layers=3):
    """Bottom + left + right walls, `layers` rows at -(l+1/2)dx outside the tank."""
    pts = []
    nb = int(round((tank_w + 2 * (layers - 0.5) * dx) / dx)) + 1
    for l in range(layers):
        y = -(l + 0.5) * dx
        xs = -(layers - 0.5) * dx + np.arange(nb) * dx
        pts.append(np.stack([xs, np.full(nb, y)], axis=1))
    ny = int(math.floor(tank_h / dx + 1e-9))
    ys = (np.arange(ny) + 0.5) * dx
    for l in range(layers):
        pts.append(np.stack([np.full(ny, -(l + 0.5) * dx), ys], axis=1))
    for l in range(layers):
        pts.append(np.stack([np.full(ny, tank_w + (l + 0.5) * dx), ys], axis=1))
    return np.concatenate(pts, axis=0)


def dam_break(width=1.0, height=2.0, dx=0.02, tank=(4.0, 3.0), layers=3, h_factor=1.3,
              alpha_av=0.1, cfl=0.25, rho0=1000.0, gamma=7.0, g=9.81, jitter=0.0, seed=0,
              eos_height=None, name=None) -> Scene:
    """Dam break: fluid column width x height at the left of a tank, hydrostatic density.

    Fluid at ((i+1/2)dx, (j+1/2)dx); walls of `layers` fixed boundary rows at the
    same spacing; m = rho0 dx^2; h = h_factor dx; c0 = 10 sqrt(2 g H);
    dt = cfl h / c0.  `jitter` displaces fluid particles uniformly by +-jitter*dx.
    `eos_height` (default: the column height) is the height c0 is derived from, so
    columns of different heights can share one sound speed and time step.
    """
    nx = int(math.floor(width / dx + 1e-9))
    ny = int(math.floor(height / dx + 1e-9))
    fl = _lattice(nx, ny, dx)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        fl = fl + rng.uniform(-jitter * dx, jitter * dx, fl.shape)
    H = ny * dx
    c0, B = tait_constants(rho0, gamma, g, H if eos_height is None else eos_height)
    rho, p = hydrostatic_density(fl[:, 1], H, rho0, gamma, g, B)
    walls = _tank_walls(dx, tank[0], tank[1], layers)
    mass = rho0 * dx * dx
    h = h_factor * dx
    k, c = _wall_params(mass, dx, g)
    margin = (layers + 2) * dx
    return _assemble(fl, rho, p, walls, np.zeros((0, 2)), mass, dx=dx, h=h, rho0=rho0,
                     gamma=gamma, c0=c0, B=B, g=g, gravity=1, alpha_av=alpha_av,
                     k_spring=k, c_damp=c, dt=cfl * h / c0,
                     box=(-margin, -margin, tank[0] + margin, tank[1] + margin),
                     name=name or f"dam_break_{width}x{height}_dx{dx:g}")


def lid_cavity(size=0.2, dx=0.005, lid_u=1.0, layers=3, h_factor=1.3, alpha_av=0.1, cfl=0.25,
               rho0=1000.0, gamma=7.0, name=None) -> Scene:
    """Lid-driven cavity: a square of fluid at rest in a closed box whose top `layers`
    rows move at lid_u (UPH_MOTION_LID, movingBoundary 3); no gravity.  The sound
    speed is 10 lid_u."""
    n = int(math.floor(size / dx + 1e-9))
    fl = _lattice(n, n, dx)
    c0 = 10.0 * lid_u
    B = rho0 * c0 * c0 / gamma
    walls = _tank_walls(dx, size, size, layers)
    nl = int(round((size + 2 * (layers - 0.5) * dx) / dx)) + 1
    lid = np.concatenate([np.stack([-(layers - 0.5) * dx + np.arange(nl) * dx,
                                    np.full(nl, size + (l + 0.5) * dx)], axis=1) for l in range(layers)])
    mass = rho0 * dx * dx
    h = h_factor * dx
    k, c = _wall_params(mass, dx, 9.81)
    margin = (layers + 2) * dx
    return _assemble(fl, np.full(len(fl), rho0), np.zeros(len(fl)), walls, lid, mass, dx=dx, h=h, rho0=rho0,
                     gamma=gamma, c0=c0, B=B, g=9.81, gravity=0, alpha_av=alpha_av, k_spring=k, c_damp=c,
                     dt=cfl * h / c0, motion=MOTION_LID, lid_u=lid_u, depth=n * dx,
                     box=(-margin, -margin, size + margin, size + margin + 1.0),
                     name=name or f"lid_cavity_{size}_dx{dx:g}")


def gate_dam_break(width=0.3, height=0.6, dx=0.01, tank=(1.2, 0.9), gate_u=1.0, layers=3, h_factor=1.3,
                   alpha_av=0.1, cfl=0.25, rho0=1000.0, gamma=7.0, g=9.81, name=None) -> Scene:
    """Dam break released by a gate: `layers` columns of moving boundary particles
    just right of the fluid column rise at gate_u (UPH_MOTION_GATE, movingBoundary 2)
    until 1.5 (height - 15 dx) / gate_u, then stop; their deck column 9 (gate_y) is
    their initial height."""
    sc = dam_break(width=width, height=height, dx=dx, tank=tank, layers=layers, h_factor=h_factor,
                   alpha_av=alpha_av, cfl=cfl, rho0=rho0, gamma=gamma, g=g)
    ny = int(math.floor(height / dx + 1e-9)) + 2
    gate = np.concatenate([np.stack([np.full(ny, width + (l + 0.5) * dx), (np.arange(ny) + 0.5) * dx], axis=1)
                           for l in range(layers)])
    nf, nfx = sc.n_fluid, sc.n_fixed
    out = _assemble(sc.pos[:nf], sc.rho[:nf], sc.pressure[:nf], sc.pos[nf:nf + nfx], gate, sc.mass[0],
                    dx=dx, h=sc.h, rho0=rho0, gamma=gamma, c0=sc.c0, B=sc.B, g=g, gravity=1,
                    alpha_av=alpha_av, k_spring=sc.k_spring, c_damp=sc.c_damp, dt=sc.dt,
                    motion=MOTION_GATE, lid_u=gate_u, depth=height, box=sc.box,
                    name=name or f"gate_dam_break_{width}x{height}_dx{dx:g}")
    out.gate_y = out.pos[:, 1].copy()
    return out


def sloshing_tank(width=0.4, depth=0.1, dx=0.005, wall_height=0.3, stroke=0.02, period=0.5, h_factor=1.3,
                  alpha_av=0.1, cfl=0.25, rho0=1000.0, gamma=7.0, g=9.81, jitter=0.0, seed=0,
                  name=None) -> Scene:
    """Sloshing tank (UPH_MOTION_SLOSH, virtualType 5): every wall moves with
    L(t) = stroke/2 (1 - cos 2 pi t / period).  The reference places each row by its
    itype alone -- the k-th moving row of itype 1 at x = k dx/2 + L, every row of
    itype 2 at x = L, of itype 3 at x = FlatBottomLength + L (inputbp.c:244-276) --
    so the moving set is one bottom row (2 nx + 1 particles at dx/2 on y = 0),
    then one left and one right column at x = 0 and x = width (dx/2 apart from
    y = dx/2 up to wall_height); the fluid lattice fills [0, width] x [0, depth]
    (`jitter`: fluid displaced uniformly by +-jitter*dx)."""
    nx = int(math.floor(width / dx + 1e-9))
    ny = int(math.floor(depth / dx + 1e-9))
    fl = _lattice(nx, ny, dx)
    if jitter > 0:
        fl = fl + np.random.default_rng(seed).uniform(-jitter * dx, jitter * dx, fl.shape)
    W, H = nx * dx, ny * dx
    c0, B = tait_constants(rho0, gamma, g, H)
    rho, p = hydrostatic_density(fl[:, 1], H, rho0, gamma, g, B)
    nb = 2 * nx + 1
    bottom = np.stack([np.arange(nb) * 0.5 * dx, np.zeros(nb)], axis=1)
    nw = int(math.floor(wall_height / (0.5 * dx) + 1e-9))
    ys = np.arange(1, nw + 1) * 0.5 * dx
    left = np.stack([np.zeros(nw), ys], axis=1)
    right = np.stack([np.full(nw, W), ys], axis=1)
    moving = np.concatenate([bottom, left, right])
    mass = rho0 * dx * dx
    h = h_factor * dx
    k, c = _wall_params(mass, dx, g)
    margin = 4 * dx + stroke
    out = _assemble(fl, rho, p, np.zeros((0, 2)), moving, mass, dx=dx, h=h, rho0=rho0, gamma=gamma,
                    c0=c0, B=B, g=g, gravity=1, alpha_av=alpha_av, k_spring=k, c_damp=c, dt=cfl * h / c0,
                    motion=MOTION_SLOSH, m_amp=stroke, m_omega=2.0 * math.pi / period, flat_bottom=W, depth=H,
                    box=(-margin, -margin, W + margin, wall_height + margin),
                    name=name or f"sloshing_{width}x{depth}_dx{dx:g}")
    nf = out.n_fluid
    out.itype[nf:nf + nb] = 1
    out.itype[nf + nb:nf + nb + nw] = 2
    out.itype[nf + nb + nw:] = 3
    out.vtype = 5          # the sloshing deck's virtualType also selects its mirror particles
    return out


def wave_tank(length=6.0, depth=1.0, dx=0.005, beach_toe=5.0, beach_slope=0.1,
              amp=0.05, period=1.2, layers=3, h_factor=1.3, alpha_av=0.1, cfl=0.25,
              rho0=1000.0, gamma=7.0, g=9.81, wall_height=1.5, jitter=0.0, seed=0, name=None) -> Scene:
    """Water layer with a piston wavemaker on the left and a sloping beach on the right.

    The piston is `layers` rows of moving boundary particles displaced by
    amp sin(2 pi t / period); the beach rises at `beach_slope` from x = beach_toe
    to the right wall at x = length.  `jitter` displaces fluid particles uniformly by
    +-jitter*dx.
    """
    nx = int(math.floor(length / dx + 1e-9))
    ny = int(math.floor(depth / dx + 1e-9))
    fl = _lattice(nx, ny, dx)
    bed = np.where(fl[:, 0] > beach_toe, (fl[:, 0] - beach_toe) * beach_slope, 0.0)
    fl = fl[fl[:, 1] >= bed + 0.5 * dx - 1e-9 * dx]     # lattice rows start at y = dx/2
    if jitter > 0:
        fl = fl + np.random.default_rng(seed).uniform(-jitter * dx, jitter * dx, fl.shape)
    H = ny * dx
    c0, B = tait_constants(rho0, gamma, g, H)
    rho, p = hydrostatic_density(fl[:, 1], H, rho0, gamma, g, B)
    fixed = []
    # flat bottom under [-(layers-1/2)dx, beach_toe]
    nbot = int(round((beach_toe + (layers - 0.5) * dx) / dx)) + 1
    for l in range(layers):
        xs = -(layers - 0.5) * dx + np.arange(nbot) * dx
        fixed.append(np.stack([xs, np.full(nbot, -(l + 0.5) * dx)], axis=1))
    # beach: rows parallel to the slope, spacing dx along it, offset normal to it
    ds = dx
    run = length - beach_toe
    L = run * math.sqrt(1 + beach_slope ** 2)
    ns = int(math.floor(L / ds)) + 1
    cth, sth = 1 / math.sqrt(1 + beach_slope ** 2), beach_slope / math.sqrt(1 + beach_slope ** 2)
    for l in range(layers):
        s = (np.arange(ns) + 1) * ds
        off = (l + 0.5) * dx
        xs = beach_toe + s * cth + off * sth
        ys = s * sth - off * cth
        fixed.append(np.stack([xs, ys], axis=1))
    # right wall above the beach top
    top = run * beach_slope
    nyr = int(math.floor((wall_height - top) / dx))
    ys = top + (np.arange(nyr) + 0.5) * dx
    for l in range(layers):
        fixed.append(np.stack([np.full(nyr, length + (l + 0.5) * dx), ys], axis=1))
    fixed = np.concatenate(fixed, axis=0)
    nyp = int(math.floor(wall_height / dx))
    ys = (np.arange(nyp) + 0.5) * dx
    piston = np.concatenate([np.stack([np.full(nyp, -(l + 0.5) * dx), ys], axis=1)
                             for l in range(layers)], axis=0)
    mass = rho0 * dx * dx
    h = h_factor * dx
    k, c = _wall_params(mass, dx, g)
    margin = (layers + 2) * dx
    sc = _assemble(fl, rho, p, fixed, piston, mass, dx=dx, h=h, rho0=rho0, gamma=gamma,
                   c0=c0, B=B, g=g, gravity=1, alpha_av=alpha_av, k_spring=k, c_damp=c,
                   dt=cfl * h / c0, motion=MOTION_SINE_OFFSET, m_amp=amp,
                   m_omega=2 * math.pi / period,
                   box=(-margin - amp, -margin, length + margin, wall_height + margin),
                   name=name or f"wave_tank_{length}x{depth}_dx{dx:g}")
    hydrostatic_walls(sc, H)
    return sc


def hydrostatic_walls(sc: Scene, H: float) -> None:
    """Give wall particles below the free surface y = H the hydrostatic density.

    Walls are reset to their deck state every step (inputbp.c:117-150), so a wall
    at rho0 carries zero pressure and supports the fluid next to it with only
    the fluid's own pressure; at the hydrostatic density it carries the pressure
    of the water column at its depth.
    """
    w = slice(sc.n_fluid, sc.n)
    y = sc.pos[w, 1]
    rho_h, p_h = hydrostatic_density(np.minimum(y, H), H, sc.rho0, sc.gamma, sc.g, sc.B)
    sc.rho[w] = np.where(y < H, rho_h, sc.rho0)
    sc.pressure[w] = np.where(y < H, p_h, 0.0)


def batch_dam_breaks(n_scenes=512, seed=20260512, width=(0.3, 1.5), height=(0.5, 2.0),
                     dx=(0.005, 0.02), max_particles: Optional[int] = None) -> List[Scene]:
    """Independent dam breaks with seeded column width, height and dx (numpy default_rng).

    `max_particles`, if given, redraws dx for scenes above that fluid count.
    """
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenes):
        w = rng.uniform(*width)
        hgt = rng.uniform(*height)
        d = rng.uniform(*dx)
        if max_particles is not None:
            while (w / d) * (hgt / d) > max_particles:
                d = rng.uniform(*dx)
        out.append(dam_break(width=w, height=hgt, dx=d, name=f"batch{s}"))
    return out


# ---------------------------------------------------------------- deck format

def write_deck(path: str, sc: Scene, which: str) -> None:
    sel = {"fluid": slice(0, sc.n_fluid),
           "fixed": slice(sc.n_fluid, sc.n_fluid + sc.n_fixed),
           "moving": slice(sc.n_fluid + sc.n_fixed, sc.n)}[which]
    idx = np.arange(sc.n)[sel]
    gate = sc.gate_y if sc.gate_y is not None else np.zeros(sc.n)
    with open(path, "w") as f:
        f.write(f"{len(idx)}\n")
        for i in idx:
            f.write("%.10f %.10f %.10f %.10f %.10f %.10f %.10f %.10f %.10f %d %.10f %.10f\n" % (
                sc.pos[i, 0], sc.pos[i, 1], sc.vel[i, 0], sc.vel[i, 1], sc.rho[i],
                sc.pressure[i], sc.mass[i], 0.0, gate[i], sc.itype[i], sc.lvel[i, 0], sc.lvel[i, 1]))


PARAMETER_NAMES = ["stroke", "wavePeriod", "LocalDepth", "FlatBottomLength", "density0",
                   "AdiabaticConstant", "Gravity", "dx", "dy", "dim", "viscosity0", "machNumber",
                   "pecletNumber", "epsilon", "alpha", "virtualType", "printStep", "screenStep",
                   "timeStep", "lidVelocity", "nearNeighbor", "selectBoundaryForce", "lennardSelect",
                   "springConstant0", "dampingConstant0", "flagInitSimul", "gravityFlag",
                   "thermalDiffusivity0", "movingBoundary", "NumberOfInitializationSteps",
                   "pressureProbeX", "pressureProbeY", "stressType", "disspationControlVolume"]


def scene_parameters(sc: Scene, **over) -> dict:
    """parameters.dat values (System.c:60-94) that reproduce a Scene's constants when
    read back by scene_from_rows: machNumber from c0 (sqrt(2 g H) / c0 for the gate,
    System.c:292-297), nearNeighbor = 4 pi (h / dx)^2, the motion as movingBoundary
    (1 piston, 3 lid, 2 gate) or virtualType 5 (sloshing)."""
    H = sc.depth if sc.depth > 0 else 1.0
    mb = {MOTION_PISTON_REF: 1, MOTION_LID: 3, MOTION_GATE: 2}.get(sc.motion, 0)
    if sc.motion not in (MOTION_NONE, MOTION_PISTON_REF, MOTION_LID, MOTION_GATE, MOTION_SLOSH):
        raise ValueError(f"motion {sc.motion} has no parameters.dat form")
    # virtualType 5 is the sloshing motion whenever movingBoundary is 0 (inputbp.c:244)
    if (sc.motion == MOTION_SLOSH) != (sc.vtype == 5 and mb == 0):
        raise ValueError(f"motion {sc.motion} with virtualType {sc.vtype} has no parameters.dat form")
    mach = math.sqrt((2.0 if mb == 2 else 1.0) * sc.g * H) / sc.c0
    prm = dict(stroke=sc.m_amp, wavePeriod=2.0 * math.pi / sc.m_omega if sc.m_omega > 0 else 1.0,
               LocalDepth=H, FlatBottomLength=sc.flat_bottom, density0=sc.rho0, AdiabaticConstant=sc.gamma,
               Gravity=sc.g, dx=sc.dx, dy=sc.dx, dim=2, viscosity0=sc.mu, machNumber=mach, pecletNumber=1,
               epsilon=sc.eps, alpha=sc.alpha_av, virtualType=sc.vtype,
               printStep=1000000, screenStep=1000000, timeStep=sc.dt, lidVelocity=sc.lid_u,
               nearNeighbor=4.0 * math.pi * (sc.h / sc.dx) ** 2, selectBoundaryForce=3, lennardSelect=0,
               springConstant0=sc.k_spring, dampingConstant0=sc.c_damp, flagInitSimul=1,
               gravityFlag=sc.gravity, thermalDiffusivity0=sc.kappa, movingBoundary=mb,
               NumberOfInitializationSteps=0, pressureProbeX=0.0, pressureProbeY=0.0,
               stressType=sc.stress_type,
               disspationControlVolume=0)
    prm.update(over)
    return prm


def write_deck_dir(d: str, sc: Scene, **over) -> None:
    """A reference input directory for a Scene: parameters.dat (scene_parameters) and
    the fluid, fixed, moving and (empty) solid particle decks."""
    prm = scene_parameters(sc, **over)
    with open(os.path.join(d, "parameters.dat"), "w") as f:
        f.write("\n".join(repr(float(prm[k])) for k in PARAMETER_NAMES) + "\n")
    write_deck(os.path.join(d, "FluidParticle.dat"), sc, "fluid")
    write_deck(os.path.join(d, "FixedBoundaryParticle.dat"), sc, "fixed")
    write_deck(os.path.join(d, "MovingBoundaryParticle.dat"), sc, "moving")
    with open(os.path.join(d, "SolidParticle.dat"), "w") as f:
        f.write("0\n")


def read_deck(path: str) -> np.ndarray:
    """Rows of a 12-column deck as an (n, 12) float64 array."""
    with open(path) as f:
        n = int(f.readline().split()[0])
        if n == 0:
            return np.zeros((0, 12))
        data = np.loadtxt(f, max_rows=n, ndmin=2)
    if data.shape[1] < 12:   # older 10-column writer: no localVelocity columns
        data = np.concatenate([data, np.zeros((n, 12 - data.shape[1]))], axis=1)
    return data[:, :12]


def read_parameters(path: str) -> dict:
    """The 34 positional values of parameters.dat (System.c:60-94)."""
    with open(path) as f:
        vals = f.read().split()
    return {k: float(v) for k, v in zip(PARAMETER_NAMES, vals)}


def deck_motion(prm: dict) -> int:
    """The moving rows' kinematics a parameters.dat selects (inputbp.c:192-283):
    movingBoundary 1, 3, 2, else virtualType 5."""
    mb = int(prm["movingBoundary"])
    if mb == 1:
        return MOTION_PISTON_REF
    if mb == 3:
        return MOTION_LID
    if mb == 2:
        return MOTION_GATE
    if int(prm["virtualType"]) == 5:
        return MOTION_SLOSH
    return MOTION_NONE


def scene_from_rows(fluid: np.ndarray, fixed: np.ndarray, moving: np.ndarray,
                    prm: dict, **over) -> Scene:
    """Build a Scene from deck rows and a parameters.dat dict.

    Derived constants follow System.c:292-313: c = sqrt(g H)/Mach (sqrt(2 g H)/Mach
    when movingBoundary == 2), B = rho0 c^2 / gamma, eta = sqrt(Nn / 4 pi),
    hsml = eta dx (input.c:107).
    """
    rows = np.concatenate([fluid, fixed, moving], axis=0)
    g = prm["Gravity"]
    H = prm["LocalDepth"]
    if int(prm["movingBoundary"]) == 2:
        c = math.sqrt(2 * g * H) / prm["machNumber"]
    else:
        c = math.sqrt(g * H) / prm["machNumber"]
    B = (1.0 / prm["AdiabaticConstant"]) * prm["density0"] * c * c
    eta = math.sqrt(prm["nearNeighbor"] / (4.0 * math.pi))
    dx = prm["dx"]
    pos = rows[:, 0:2].copy()
    lo = pos.min(axis=0) - 5 * dx
    hi = pos.max(axis=0) + 5 * dx
    kw = dict(dx=dx, h=eta * dx, rho0=prm["density0"], gamma=prm["AdiabaticConstant"],
              c0=c, B=B, g=g, gravity=int(prm["gravityFlag"]), alpha_av=prm["alpha"],
              k_spring=prm["springConstant0"], c_damp=prm["dampingConstant0"],
              dt=prm["timeStep"], motion=deck_motion(prm),
              lid_u=prm["lidVelocity"], depth=prm["LocalDepth"], flat_bottom=prm["FlatBottomLength"],
              m_amp=prm["stroke"], m_omega=2.0 * math.pi / prm["wavePeriod"],
              kappa=prm["thermalDiffusivity0"], mu=prm["viscosity0"], eps=prm["epsilon"],
              vtype=int(prm["virtualType"]), stress_type=int(prm["stressType"]),
              box=(lo[0], lo[1], hi[0], hi[1]), name="deck")
    kw.update(over)
    return Scene(pos=pos, vel=rows[:, 2:4].copy(), rho=rows[:, 4].copy(),
                 pressure=rows[:, 5].copy(), mass=rows[:, 6].copy(), gate_y=rows[:, 8].copy(),
                 itype=rows[:, 9].astype(np.int32), lvel=rows[:, 10:12].copy(),
                 n_fluid=len(fluid), n_fixed=len(fixed), n_moving=len(moving), **kw)


def load_deck_dir(d: str, fluid_file="FluidParticle.dat", **over) -> Scene:
    prm = read_parameters(os.path.join(d, "parameters.dat"))
    fl = read_deck(os.path.join(d, fluid_file))
    fx = read_deck(os.path.join(d, "FixedBoundaryParticle.dat"))
    mv = read_deck(os.path.join(d, "MovingBoundaryParticle.dat"))
    return scene_from_rows(fl, fx, mv, prm, **over)
