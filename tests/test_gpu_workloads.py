"""The wave tank, the 512-scene batch, the autotuner and the C host driver, on the GPU."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_golden_scene
from uph_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-workloads] {msg}", flush=True)


def test_wave_tank_240k_hydrostatic_profile():
    """6.0 x 1.0 m layer at dx = 5 mm (~240K fluid), sinusoidal piston, 1:10 beach, 5000 steps."""
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    sc = scenes.wave_tank()
    _say(f"wave tank: {sc.n_fluid} fluid + {sc.n_boundary} wall, 300 steps vs oracle")
    g = GpuSolver(sc)
    g.step(300)
    o = Sim(sc, nthreads=0)
    o.step(300)
    nf = sc.n_fluid
    sg, so = g.state(), o.state()
    drho = float(np.max(np.abs(sg["rho"][:nf] - so["rho"][:nf]) / so["rho"][:nf]))
    dxe = float(np.abs(sg["x"][:nf] - so["x"][:nf]).max()) / sc.dx
    mv = slice(sc.n_fluid + sc.n_fixed, sc.n)
    dpist = float(np.abs(sg["x"][mv] - so["x"][mv]).max())
    _say(f"wave tank 300 steps: max rel rho err {drho:.2e}, max |dx|/dx {dxe:.2e}, piston {dpist:.1e} m")
    assert drho < 1e-5
    assert dxe < 1e-3
    assert dpist < 1e-12
    del o, so
    # 5000 steps; the hydrostatic profile is checked on its average over the last 2500
    # steps (about one period of the column's acoustic breathing mode, 4 H / c0)
    g.step(2200)
    H = 1.0
    bins = np.linspace(0.05, 0.95, 7)
    acc = []
    for _ in range(25):
        g.step(100)
        st = g.state()
        x, y, rho = st["x"][:nf, 0], st["x"][:nf, 1], st["rho"][:nf]
        assert np.all(np.isfinite(st["x"])) and np.all(np.isfinite(rho))
        P = sc.B * ((rho / sc.rho0) ** 7 - 1.0)
        m = (x > 2.0) & (x < 4.0)
        prof = []
        for lo, hi in zip(bins[:-1], bins[1:]):
            sel = m & (H - y > lo) & (H - y <= hi)
            prof.append(P[sel].mean() / (sc.rho0 * sc.g * (H - y[sel]).mean()))
        acc.append(prof)
    assert g.sync() == 0
    mean = np.mean(acc, axis=0)
    _say(f"wave tank P / (rho0 g depth) by depth, mean over steps 2600-5000: {np.round(mean, 4)}; "
         f"instantaneous range [{np.min(acc):.3f}, {np.max(acc):.3f}]")
    assert np.all(np.abs(mean - 1.0) < 0.03)


def test_batch_of_512_dam_breaks():
    """512 dam breaks with seeded width/height/dx in one batch: every scene equals its own
    single-scene run, and two small scenes match the oracle."""
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    scs = scenes.batch_dam_breaks()
    nf = [s.n_fluid for s in scs]
    _say(f"batch: {len(scs)} scenes, {sum(nf)} fluid (min {min(nf)}, max {max(nf)})")
    gb = GpuSolver(scs, use_graph=True)
    gb.step(20)
    assert gb.sync() == 0
    rng = np.random.default_rng(0)
    for k in rng.choice(len(scs), 4, replace=False):
        gs = GpuSolver(scs[k], sort_every_step=True)   # the captured step sorts every step
        gs.step(20)
        a, b = gb.state(int(k)), gs.state(0)
        for key in ("x", "v", "rho"):
            assert np.array_equal(a[key], b[key]), (k, key)
    small = np.argsort(nf)[:2]
    for k in small:
        o = Sim(scs[k])
        o.step(20)
        so, sg = o.state(), gb.state(int(k))
        n = scs[k].n_fluid
        drho = float(np.max(np.abs(sg["rho"][:n] - so["rho"][:n]) / so["rho"][:n]))
        assert drho < 1e-6, (k, drho)
    tot, _ = gb.time_steps(20)
    _say(f"batch: {tot / 20:.3f} ms per batch step, "
         f"{sum(nf) * 20 / (tot * 1e-3):.3e} particle-updates/s")
    gb.step(460)                       # 500 steps in all
    assert gb.sync() == 0
    for k in (0, 100, 511):
        st = gb.state(k)
        n = scs[k].n_fluid
        assert np.all(np.isfinite(st["x"])) and 950 < st["rho"][:n].min() <= st["rho"][:n].max() < 1100


def test_batch_of_512_dam_breaks_lazy_resort():
    """The same batch with the lazy re-sort asked for (the default sorts a batch of this
    size every step), eager launches: 300 steps with re-sorts; sampled scenes equal
    their own single-scene lazy runs with the same skin bit for bit (each scene's
    re-sorts depend on its own particles only), the smallest matches the oracle."""
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    scs = scenes.batch_dam_breaks()
    _say("lazy batch of 512: 300 steps")
    gb = GpuSolver(scs, sort_every_step=False, skin=0.05)
    assert gb.skin() == 0.05
    gb.step(300)
    assert gb.sync() == 0
    nres = gb.resorts()
    _say(f"lazy batch: {nres} re-sorts in 300 steps")
    assert nres >= 1
    rng = np.random.default_rng(1)
    for k in [int(np.argmin([s.n_fluid for s in scs]))] + [int(j) for j in rng.choice(len(scs), 3, replace=False)]:
        gs = GpuSolver(scs[k], skin=0.05)
        gs.step(300)
        a, b = gb.state(k), gs.state(0)
        for key in ("x", "v", "rho"):
            assert np.array_equal(a[key], b[key]), (k, key)
    k = int(np.argmin([s.n_fluid for s in scs]))
    o = Sim(scs[k])
    o.step(300)
    so, sg = o.state(), gb.state(k)
    n = scs[k].n_fluid
    drho = float(np.max(np.abs(sg["rho"][:n] - so["rho"][:n]) / so["rho"][:n]))
    _say(f"lazy batch: smallest scene ({n} fluid) vs oracle after 300 steps: rho {drho:.2e}")
    assert drho < 1e-6, drho


def test_autotune_on_gpu():
    from oracle.oracle import Sim
    from uph_amd import autotune
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(dx=0.004)
    res = autotune.autotune(sc, warmup=2, steps=5, log=_say)
    best = res["best"]
    _say(f"autotune best {best} at {res['ms_per_step']:.4f} ms/step")
    g = GpuSolver(sc, **best)
    g.step(50)
    o = Sim(sc, nthreads=0)
    o.step(50)
    n = sc.n_fluid
    sg, so = g.state(), o.state()
    assert float(np.max(np.abs(sg["rho"][:n] - so["rho"][:n]) / so["rho"][:n])) < 1e-6


@pytest.mark.parametrize("cf,blk", [(1.3, 128), (1.15, 512)])
def test_cell_factor_and_block_match_oracle(cf, blk):
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(jitter=0.1, seed=3)
    g = GpuSolver(sc, cell_factor=cf, pair_block=blk)
    g.step(200)
    o = Sim(sc, nthreads=0)
    o.step(200)
    n = sc.n_fluid
    sg, so = g.state(), o.state()
    assert float(np.max(np.abs(sg["rho"][:n] - so["rho"][:n]) / so["rho"][:n])) < 1e-6


def test_c_host_driver(tmp_path):
    """uph_run (C, links only the C ABI) on the reference hydrostatic deck."""
    from uph_amd import solver
    from uph_amd.scenes import write_deck
    drv = os.path.join(ROOT, "uph_amd", "uph_run")
    assert os.path.exists(drv), "uph_run not built"
    z = np.load(os.path.join(GOLDEN, "hydrostatic_init.npz"), allow_pickle=False)
    vals = z["param_values"]
    sc = load_golden_scene("hydrostatic_init.npz", alpha_av=0.1)
    with open(tmp_path / "parameters.dat", "w") as f:
        f.write("\n".join(repr(float(v)) if not float(v).is_integer() else str(int(v))
                          for v in vals) + "\n")
    write_deck(str(tmp_path / "FluidParticle.dat"), sc, "fluid")
    write_deck(str(tmp_path / "FixedBoundaryParticle.dat"), sc, "fixed")
    write_deck(str(tmp_path / "MovingBoundaryParticle.dat"), sc, "moving")
    (tmp_path / "SolidParticle.dat").write_text("0\n")
    _say("uph_run 100 steps")
    out = subprocess.run([drv, "100", "--alpha", "0.1"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "particle-updates/s" in out.stdout
    rows = np.loadtxt(tmp_path / "FluidParticleFinal.dat", skiprows=1)
    # the same run through the Python mirror, on the same decks and the driver's cell box
    from uph_amd.scenes import load_deck_dir
    sc2 = load_deck_dir(str(tmp_path), alpha_av=0.1)
    lo, hi = sc2.pos.min(0), sc2.pos.max(0)
    mg = 10 * sc2.h
    sc2.box = (lo[0] - mg, lo[1] - mg, hi[0] + mg, hi[1] + 2 * (hi[1] - lo[1]))
    g = solver.GpuSolver(sc2)
    g.step(100)
    st = g.state()
    n = sc2.n_fluid
    ref = np.loadtxt(tmp_path / "FluidParticle.dat", skiprows=1)
    assert not np.allclose(rows[:, 0:2], ref[:, 0:2], rtol=0, atol=1e-12)   # it moved
    # the driver prints with 10 decimals
    assert np.allclose(rows[:, 0:2], st["x"][:n], rtol=0, atol=6e-11)
    assert np.allclose(rows[:, 4], st["rho"][:n], rtol=0, atol=6e-11)


def _deck_dir(tmp_path, fixture, print_step=None):
    from uph_amd.scenes import write_deck
    z = np.load(os.path.join(GOLDEN, fixture), allow_pickle=False)
    vals = z["param_values"].copy()
    if print_step is not None:
        vals[16] = print_step           # printStep (System.c:60-94 field 17)
    sc = load_golden_scene(fixture, alpha_av=0.1)
    with open(tmp_path / "parameters.dat", "w") as f:
        f.write("\n".join(repr(float(v)) if not float(v).is_integer() else str(int(v))
                          for v in vals) + "\n")
    write_deck(str(tmp_path / "FluidParticle.dat"), sc, "fluid")
    write_deck(str(tmp_path / "FixedBoundaryParticle.dat"), sc, "fixed")
    write_deck(str(tmp_path / "MovingBoundaryParticle.dat"), sc, "moving")
    (tmp_path / "SolidParticle.dat").write_text("0\n")
    return sc


@pytest.mark.parametrize("fixture", ["hydrostatic_init.npz", "waves.npz"])
def test_c_host_driver_compat_derivatives(tmp_path, fixture):
    """The reference's host loop (sph.cc:88-168) in C over a System object, calling the
    library's derivatives() each step, against the device-resident loop and the oracle."""
    from oracle.oracle import Sim
    from uph_amd.scenes import load_deck_dir
    drv = os.path.join(ROOT, "uph_amd", "uph_run")
    assert os.path.exists(drv), "uph_run not built"
    _deck_dir(tmp_path, fixture)
    steps = 100
    _say(f"uph_run --compat {steps} on {fixture}")
    out = subprocess.run([drv, str(steps), "--alpha", "0.1", "--compat", "--out", "compat.dat"],
                         cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    _say(out.stdout.strip())
    diag = [l for l in out.stdout.splitlines() if l.startswith("diagnostics:")]
    assert diag and "nan" not in diag[0].lower()
    out = subprocess.run([drv, str(steps), "--alpha", "0.1", "--out", "device.dat"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    a = np.loadtxt(tmp_path / "compat.dat", skiprows=1)
    b = np.loadtxt(tmp_path / "device.dat", skiprows=1)
    sc = load_deck_dir(str(tmp_path), alpha_av=0.1)
    dx = sc.dx
    dpos = float(np.abs(a[:, 0:2] - b[:, 0:2]).max()) / dx
    drho = float(np.max(np.abs(a[:, 4] - b[:, 4]) / b[:, 4]))
    _say(f"compat vs device loop: max |dx|/dx {dpos:.2e}, max rel rho {drho:.2e}")
    # same arithmetic (the System receives the FP64 acceleration the device corrector
    # uses); the compat context picks its own cell box, so FP32 sums may run in
    # another order
    assert dpos < 1e-8 and drho < 1e-9
    # and against the FP64 oracle on the same decks
    o = Sim(sc, nthreads=0)
    o.step(steps)
    so = o.state()
    n = sc.n_fluid
    drho_o = float(np.max(np.abs(a[:, 4] - so["rho"][:n]) / so["rho"][:n]))
    dpos_o = float(np.abs(a[:, 0:2] - so["x"][:n]).max()) / dx
    _say(f"compat vs oracle: max rel rho {drho_o:.2e}, max |dx|/dx {dpos_o:.2e}")
    assert drho_o < 1e-6 and dpos_o < 1e-4


def test_c_host_driver_writes_the_reference_output(tmp_path):
    """printStep output (output.c:19-130) from both driver loops: out<itime>.dat with the
    reference's 14 columns for every particle, one line per output in
    TotalKineticEnergy.dat and PointPressure.dat; the two loops' snapshots agree."""
    drv = os.path.join(ROOT, "uph_amd", "uph_run")
    sc = _deck_dir(tmp_path, "hydrostatic_init.npz", print_step=25)
    for mode in ("device", "compat"):
        d = tmp_path / mode
        d.mkdir()
        for f in ("parameters.dat", "FluidParticle.dat", "FixedBoundaryParticle.dat",
                  "MovingBoundaryParticle.dat", "SolidParticle.dat"):
            (d / f).write_text((tmp_path / f).read_text())
        args = [drv, "100", "--alpha", "0.1"] + (["--compat"] if mode == "compat" else [])
        out = subprocess.run(args, cwd=d, capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr
        for it in (25, 50, 75, 100):
            rows = np.loadtxt(d / f"out{it}.dat")
            assert rows.shape == (sc.n, 14), rows.shape
            assert np.array_equal(rows[:, 0], np.arange(sc.n))
        ke = np.loadtxt(d / "TotalKineticEnergy.dat")
        pp = np.loadtxt(d / "PointPressure.dat")
        p2 = np.loadtxt(d / "PointPressureTwo.dat")
        assert ke.shape == (4, 3) and pp.shape == (4, 5) and p2.shape == (4, 3)
        assert np.allclose(ke[:, 0], np.array([25, 50, 75, 100]) * sc.dt, rtol=1e-9)
        # the second probe (at the first probe's location by default): its mean of the
        # positive pressures within 2h, in both files
        assert np.all(p2[:, 1] > 0) and np.array_equal(pp[:, 4], p2[:, 2])
        # damGateFlag (input.c:148-159): fluid rows by their deck x, 1 up to 0.38 m, else -1
        rows = np.loadtxt(d / "out100.dat")
        flag = np.where(sc.pos[:sc.n_fluid, 0] <= 0.38, 1, -1)
        assert np.array_equal(rows[:sc.n_fluid, 9], flag) and np.all(rows[sc.n_fluid:, 9] == 0)
        assert np.any(rows[:sc.n_fluid, 10] != 0) and np.all(rows[sc.n_fluid:, 10] == 0)
        final = np.loadtxt(d / "FluidParticleFinal.dat", skiprows=1)
        snap = np.loadtxt(d / "out100.dat")[:sc.n_fluid]
        assert np.allclose(snap[:, 1:3], final[:, 0:2], rtol=0, atol=2e-10)
    a = np.loadtxt(tmp_path / "device" / "out100.dat")
    b = np.loadtxt(tmp_path / "compat" / "out100.dat")
    nf = sc.n_fluid
    _say(f"out100.dat device vs compat: max |dx| {np.abs(a[:nf, 1:3] - b[:nf, 1:3]).max():.2e}")
    assert np.allclose(a[:nf, 1:6], b[:nf, 1:6], rtol=0, atol=1e-9)
    # approxVorticity of the last evaluation: the device loop's pair kernel inputs and the
    # state derivatives() was called with are the same state
    wa, wb = a[:nf, 10], b[:nf, 10]
    _say(f"approxVorticity device vs compat: max |diff| {np.abs(wa - wb).max():.2e} of max {np.abs(wb).max():.2e}")
    assert np.abs(wa - wb).max() <= 1e-6 * np.abs(wb).max()


def test_c_host_driver_restart_from_fluid_particle_init(tmp_path):
    """FluidParticleInit.dat (output.c:113-128), written at step NumberOfInitializationSteps,
    holds the fluid rows of that step in the deck format: its values are the snapshot's at
    the printed decimals, and a run started from it (as FluidParticle.dat) by the C driver
    equals the Python host started from the same file, bit for bit at the printed
    decimals.  Against the uninterrupted run the restart differs by the rounding to 10
    decimals and the first-step form of the leapfrog (sph.cc:154-167), which it takes
    again (half a step of the rates where the run would take a whole one): reported,
    and bounded loosely."""
    import shutil
    from uph_amd.scenes import load_deck_dir
    from uph_amd.solver import GpuSolver
    drv = os.path.join(ROOT, "uph_amd", "uph_run")
    sc = _deck_dir(tmp_path, "hydrostatic_init.npz", print_step=25)
    prm = (tmp_path / "parameters.dat").read_text().split()
    prm[29] = "50"                          # NumberOfInitializationSteps (field 30)
    (tmp_path / "parameters.dat").write_text("\n".join(prm) + "\n")
    out = subprocess.run([drv, "100", "--alpha", "0.1"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    init = np.loadtxt(tmp_path / "FluidParticleInit.dat", skiprows=1)
    with open(tmp_path / "FluidParticleInit.dat") as f:
        assert int(f.readline()) == sc.n_fluid
    snap = np.loadtxt(tmp_path / "out50.dat")[:sc.n_fluid]
    assert init.shape == (sc.n_fluid, 12)
    assert np.allclose(init[:, 0:5], snap[:, 1:6], rtol=0, atol=1.5e-10)     # x, y, vx, vy, rho
    assert np.allclose(init[:, 5], snap[:, 8], rtol=0, atol=1.5e-10)         # DynamicPressure
    assert np.array_equal(init[:, 9], np.ones(sc.n_fluid))                 # itype
    rd = tmp_path / "restart"
    rd.mkdir()
    for f in ("parameters.dat", "FixedBoundaryParticle.dat", "MovingBoundaryParticle.dat", "SolidParticle.dat"):
        shutil.copy(tmp_path / f, rd / f)
    shutil.copy(tmp_path / "FluidParticleInit.dat", rd / "FluidParticle.dat")
    out = subprocess.run([drv, "50", "--alpha", "0.1", "--no-output"], cwd=rd, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    c_fin = np.loadtxt(rd / "FluidParticleFinal.dat", skiprows=1)
    rs = load_deck_dir(str(rd), alpha_av=0.1)
    g = GpuSolver(rs)
    g.step(50)
    st = g.state()
    nf = sc.n_fluid
    py_fin = np.column_stack([st["x"][:nf], st["v"][:nf], st["rho"][:nf]])
    printed = np.array([float("%.10f" % v) for v in py_fin.ravel()]).reshape(py_fin.shape)
    assert np.array_equal(printed, c_fin[:, 0:5])
    full = np.loadtxt(tmp_path / "FluidParticleFinal.dat", skiprows=1)
    drho = float(np.max(np.abs(c_fin[:, 4] - full[:, 4]) / full[:, 4]))
    dx = float(np.abs(c_fin[:, 0:2] - full[:, 0:2]).max()) / sc.dx
    _say(f"restart at step 50 vs the uninterrupted run at step 100: rho {drho:.2e}, x {dx:.2e} dx")
    assert drho < 1e-3 and dx < 1e-2


def test_c_host_driver_poly7_on_the_waves_deck(tmp_path):
    """uph_run --kernel poly7: the reference's own smoothing kernel (derivatives.c:767-823),
    its waves deck, hsml = eta dx and piston, against the FP64 oracle with that kernel."""
    from oracle.oracle import Sim
    from uph_amd import scenes as S
    from uph_amd.scenes import load_deck_dir
    drv = os.path.join(ROOT, "uph_amd", "uph_run")
    assert os.path.exists(drv), "uph_run not built"
    _deck_dir(tmp_path, "waves.npz")
    steps = 100
    _say(f"uph_run --kernel poly7 {steps} on waves.npz")
    out = subprocess.run([drv, str(steps), "--alpha", "0.1", "--kernel", "poly7", "--no-output",
                          "--out", "poly7.dat"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    a = np.loadtxt(tmp_path / "poly7.dat", skiprows=1)
    sc = load_deck_dir(str(tmp_path), alpha_av=0.1)
    sc.kernel = S.KERNEL_POLY7
    o = Sim(sc, nthreads=0)
    o.step(steps)
    so = o.state()
    n = sc.n_fluid
    drho = float(np.max(np.abs(a[:, 4] - so["rho"][:n]) / so["rho"][:n]))
    dpos = float(np.abs(a[:, 0:2] - so["x"][:n]).max()) / sc.dx
    _say(f"poly7 driver vs oracle: max rel rho {drho:.2e}, max |dx|/dx {dpos:.2e}")
    assert drho < 1e-6 and dpos < 1e-4
    bad = subprocess.run([drv, "1", "--kernel", "poly9"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0 and "unknown kernel" in bad.stderr


def test_c_host_driver_reference_scheme_on_its_dambreak_deck(tmp_path):
    """uph_run --scheme ref --kernel poly7: the reference's own configuration of its
    dambreak deck (PolyOrderSeven at h = 2.691 dx with the 40-neighbour cap binding,
    pressure transport from the deck's DynamicPressure column, viscosity0, epsilon,
    thermalDiffusivity0, springs) against the oracle; the written DynamicPressure
    column is the integrated state."""
    from oracle.oracle import Sim
    from uph_amd import scenes as S
    from uph_amd.scenes import load_deck_dir
    drv = os.path.join(ROOT, "uph_amd", "uph_run")
    assert os.path.exists(drv), "uph_run not built"
    _deck_dir(tmp_path, "dambreak.npz")
    steps = 60
    _say(f"uph_run --scheme ref --kernel poly7 {steps} on dambreak.npz")
    out = subprocess.run([drv, str(steps), "--scheme", "ref", "--kernel", "poly7", "--no-output",
                          "--out", "ref.dat"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    _say(out.stdout.strip())
    a = np.loadtxt(tmp_path / "ref.dat", skiprows=1)
    sc = load_deck_dir(str(tmp_path), mode=S.MODE_REF, kernel=S.KERNEL_POLY7)
    o = Sim(sc, nthreads=0)
    o.step(steps)
    so, Po = o.state(), o.ref()["P"]
    n = sc.n_fluid
    drho = float(np.max(np.abs(a[:, 4] - so["rho"][:n]) / so["rho"][:n]))
    dP = float(np.abs(a[:, 5] - Po[:n]).max() / np.abs(Po[:n]).max())
    dpos = float(np.abs(a[:, 0:2] - so["x"][:n]).max()) / sc.dx
    _say(f"reference scheme driver vs oracle: rho {drho:.2e}, P rel-to-max {dP:.2e}, |dx|/dx {dpos:.2e}")
    assert drho < 1e-6 and dP < 1e-5 and dpos < 1e-4
    bad = subprocess.run([drv, "1", "--scheme", "reference"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0 and "unknown scheme" in bad.stderr


def test_c_host_driver_compat_reference_scheme(tmp_path):
    """The reference's host loop (sph.cc:88-168, with its fluctuation terms) over a
    System object, calling the library's derivatives() with the reference's own scheme
    and kernel (uph_derivatives_scheme(UPH_MODE_REF, UPH_KERNEL_POLY7)), on the
    reference's dambreak deck: against the device-resident loop and the oracle."""
    from oracle.oracle import Sim
    from uph_amd import scenes as S
    from uph_amd.scenes import load_deck_dir
    drv = os.path.join(ROOT, "uph_amd", "uph_run")
    assert os.path.exists(drv), "uph_run not built"
    _deck_dir(tmp_path, "dambreak.npz")
    steps = 40
    args = [drv, str(steps), "--scheme", "ref", "--kernel", "poly7", "--no-output"]
    _say(f"uph_run --compat --scheme ref --kernel poly7 {steps} on dambreak.npz")
    out = subprocess.run(args + ["--compat", "--out", "compat.dat"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    _say(out.stdout.strip())
    out = subprocess.run(args + ["--out", "device.dat"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    a = np.loadtxt(tmp_path / "compat.dat", skiprows=1)
    b = np.loadtxt(tmp_path / "device.dat", skiprows=1)
    sc = load_deck_dir(str(tmp_path), mode=S.MODE_REF, kernel=S.KERNEL_POLY7)
    dx = sc.dx
    dpos = float(np.abs(a[:, 0:2] - b[:, 0:2]).max()) / dx
    drho = float(np.max(np.abs(a[:, 4] - b[:, 4]) / b[:, 4]))
    dP = float(np.abs(a[:, 5] - b[:, 5]).max() / np.abs(b[:, 5]).max())
    _say(f"compat vs device loop: |dx|/dx {dpos:.2e}, rho {drho:.2e}, P {dP:.2e}")
    assert dpos < 1e-7 and drho < 1e-8 and dP < 1e-6
    o = Sim(sc, nthreads=0)
    o.step(steps)
    so, Po = o.state(), o.ref()["P"]
    n = sc.n_fluid
    drho_o = float(np.max(np.abs(a[:, 4] - so["rho"][:n]) / so["rho"][:n]))
    dP_o = float(np.abs(a[:, 5] - Po[:n]).max() / np.abs(Po[:n]).max())
    _say(f"compat vs oracle: rho {drho_o:.2e}, P {dP_o:.2e}")
    assert drho_o < 1e-6 and dP_o < 1e-5


def _motion_scene(motion):
    if motion == "lid":
        return scenes.lid_cavity(size=0.15, dx=0.005, lid_u=1.0)
    if motion == "gate":
        sc = scenes.gate_dam_break(width=0.3, height=0.6, dx=0.01, gate_u=2.0)
        sc.depth = 15 * sc.dx + 2.0 * 60.5 * sc.dt / 1.5      # the gate stops after 60 steps
        return sc
    return scenes.sloshing_tank(width=0.2, depth=0.06, dx=0.004, wall_height=0.12, stroke=0.01, period=0.1)


@pytest.mark.parametrize("motion", ["lid", "gate", "slosh"])
def test_c_host_driver_moving_boundaries(tmp_path, motion):
    """A reference input directory with movingBoundary 3 (lid), 2 (gate, heights in
    deck column 9) or virtualType 5 (sloshing tank): the device loop and the host loop
    over derivatives(System*) against each other and the oracle, fluid and walls."""
    from oracle.oracle import Sim
    drv = os.path.join(ROOT, "uph_amd", "uph_run")
    assert os.path.exists(drv), "uph_run not built"
    sc = _motion_scene(motion)
    scenes.write_deck_dir(str(tmp_path), sc, printStep=120)
    steps = 120
    for mode in ("device", "compat"):
        d = tmp_path / mode
        d.mkdir()
        for f in ("parameters.dat", "FluidParticle.dat", "FixedBoundaryParticle.dat",
                  "MovingBoundaryParticle.dat", "SolidParticle.dat"):
            (d / f).write_text((tmp_path / f).read_text())
        _say(f"uph_run {mode} {steps} steps, {motion} deck, n={sc.n}")
        out = subprocess.run([drv, str(steps)] + (["--compat"] if mode == "compat" else []), cwd=d,
                             capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr
    a = np.loadtxt(tmp_path / "device" / f"out{steps}.dat")
    b = np.loadtxt(tmp_path / "compat" / f"out{steps}.dat")
    back = scenes.load_deck_dir(str(tmp_path))
    assert back.motion == sc.motion
    nf, dx = back.n_fluid, back.dx
    dpos = float(np.abs(a[:, 1:3] - b[:, 1:3]).max()) / dx
    drho = float(np.max(np.abs(a[:nf, 5] - b[:nf, 5]) / b[:nf, 5]))
    _say(f"{motion}: device vs compat |dx|/dx {dpos:.2e}, rho {drho:.2e}")
    assert dpos < 1e-8 and drho < 1e-9
    o = Sim(back, nthreads=0)
    o.step(steps)
    so, go = o.state(), o.ghosts()
    n = back.n
    # the sloshing deck's virtualType 5 lists mirror rows after the particle rows
    assert b.shape[0] == n + len(go["x"])
    drho_o = float(np.max(np.abs(b[:nf, 5] - so["rho"][:nf]) / so["rho"][:nf]))
    dpos_o = float(np.abs(b[:nf, 1:3] - so["x"][:nf]).max()) / dx
    # the snapshot's boundary rows are where the last derivatives() call put them
    dwall = float(np.abs(b[nf:n, 1:3] - so["x"][nf:]).max())
    _say(f"{motion}: compat vs oracle rho {drho_o:.2e}, |dx|/dx {dpos_o:.2e}, walls {dwall:.2e} m, "
         f"{len(go['x'])} mirror rows")
    assert drho_o < 1e-6 and dpos_o < 1e-4 and dwall < 2e-10
    if len(go["x"]):
        assert float(np.abs(b[n:, 1:3] - go["x"]).max()) / dx < 1e-4


def test_c_host_driver_mirror_particles(tmp_path):
    """A deck with virtualType 8 where K binds (h = 2.691 dx): both driver loops rank the
    mirror rows in the search and list them after the particle rows of out<itime>.dat
    (output.c:28, itype 4); against each other and the oracle."""
    from oracle.oracle import Sim
    drv = os.path.join(ROOT, "uph_amd", "uph_run")
    W, H, dx = 0.24, 0.2, 0.01
    sc = scenes.dam_break(width=W, height=H, dx=dx, tank=(W, 2 * H), h_factor=2.691, jitter=0.1, seed=3)
    sc.flat_bottom, sc.depth, sc.vtype = W, H, 8
    steps = 40
    scenes.write_deck_dir(str(tmp_path), sc, printStep=steps)
    for mode in ("device", "compat"):
        d = tmp_path / mode
        d.mkdir()
        for f in ("parameters.dat", "FluidParticle.dat", "FixedBoundaryParticle.dat",
                  "MovingBoundaryParticle.dat", "SolidParticle.dat"):
            (d / f).write_text((tmp_path / f).read_text())
        _say(f"uph_run {mode} {steps} steps, virtualType 8 deck, n={sc.n}")
        out = subprocess.run([drv, str(steps)] + (["--compat"] if mode == "compat" else []), cwd=d,
                             capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr
    a = np.loadtxt(tmp_path / "device" / f"out{steps}.dat")
    b = np.loadtxt(tmp_path / "compat" / f"out{steps}.dat")
    back = scenes.load_deck_dir(str(tmp_path))
    assert back.vtype == 8
    o = Sim(back, nthreads=0)
    o.step(steps)
    so, go = o.state(), o.ghosts()
    n, nf = back.n, back.n_fluid
    ng = len(go["x"])
    assert ng > 0 and a.shape[0] == b.shape[0] == n + ng, (a.shape, b.shape, n, ng)
    assert np.all(a[n:, 11] == 4) and np.all(b[n:, 11] == 4)
    dpos = float(np.abs(a[:, 1:3] - b[:, 1:3]).max()) / dx
    drho = float(np.max(np.abs(a[:, 5] - b[:, 5]) / b[:, 5]))
    _say(f"device vs compat, particle and mirror rows: |dx|/dx {dpos:.2e}, rho {drho:.2e}")
    assert dpos < 1e-8 and drho < 1e-9
    drho_o = float(np.max(np.abs(b[:nf, 5] - so["rho"][:nf]) / so["rho"][:nf]))
    dpos_o = float(np.abs(b[:nf, 1:3] - so["x"][:nf]).max()) / dx
    dg_o = float(np.abs(b[n:, 1:3] - go["x"]).max()) / dx
    _say(f"compat vs oracle: rho {drho_o:.2e}, |dx|/dx {dpos_o:.2e}, mirror rows |dx|/dx {dg_o:.2e}")
    assert drho_o < 1e-6 and dpos_o < 1e-4 and dg_o < 1e-4
