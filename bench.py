"""Benchmark: particle-updates/s of the plain SPH step on the 1M-particle dam break.

Workload: dam break, fluid column 1.0 m x 2.0 m at dx = sqrt(2) mm (999,698 fluid
particles) in a 4 m x 3 m tank with 3 layers of fixed wall particles (21,228),
cubic spline h = 1.3 dx, Tait EOS, artificial viscosity 0.1, dt = 0.25 h / c0.
A step is one full leapfrog step of every particle on one GPU: cell binning and
sort, fused neighbour search + pair sums, FP64 integration.  Inputs are resident
in HBM before the timed region; the timed region is K steps bracketed by stream
synchronisation, timed with two HIP events on the library's stream around the
loop.  K further steps with events between the phases give the split: sort
kernels, the fused pair kernel (the dominant kernel, used for the roofline) and
the capped pass.

cpu_baseline: the FP64 CPU oracle (C/OpenMP restatement of the reference's
path, oracle/uph_oracle.c) on the same workload for a bounded number of steps.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = 80.0     # FP64 x, y, vx, vy, rho read once and written once


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    """(physical cores, logical CPUs) of this host."""
    phys = set()
    try:
        with open("/proc/cpuinfo") as f:
            pid = cid = None
            for line in f:
                if line.startswith("physical id"):
                    pid = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    cid = line.split(":")[1].strip()
                elif not line.strip():
                    if cid is not None:
                        phys.add((pid, cid))
                    pid = cid = None
    except OSError:
        pass
    return (len(phys) or None), os.cpu_count()


def cpu_threads():
    """Threads for the CPU baseline: the host-thread share this process is granted
    (OMP_NUM_THREADS when set, else the CPUs of its affinity mask)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(sc, budget_s):
    from oracle.oracle import Sim
    nthreads = cpu_threads()
    phys, logical = host_cores()
    o = Sim(sc, nthreads=nthreads)
    t0 = time.perf_counter()
    o.step(1)
    t1 = time.perf_counter()
    first = t1 - t0
    steps = max(1, min(50, int(budget_s / max(first, 1e-3))))
    t0 = time.perf_counter()
    o.step(steps)
    el = time.perf_counter() - t0
    return dict(value=sc.n_fluid * steps / el, unit="particle-updates/s", cores=nthreads,
                kind="port", host_physical_cores=phys, host_logical_cpus=logical,
                sample=f"{steps} steps of the same 1M dam break after 1 untimed step, "
                       f"FP64 oracle (C/OpenMP, {nthreads} threads: the host-thread share granted to "
                       f"this process (OMP_NUM_THREADS); host has {phys} physical cores, "
                       f"{logical} logical CPUs), {el:.1f} s")


def traffic_from_profiles():
    p = os.path.join(ROOT, "profiles", "pair_kernel_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def copy_probe(mib):
    """One device-to-device hipMemcpy of `mib` MiB: the reference dispatch for scaling the
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = int(mib) << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": n}), flush=True)


def compare_launch(steps: int, warmup: int):
    """ms per step of the ways of launching the same steps: one launch per kernel
    (uph_step), the step captured as a hipGraph and replayed, and all steps in one
    persistent cooperative kernel (uph_step_persistent), each re-sorting every step; and
    the lazy re-sort (the default: the pair kernel and one cooperative launch per step);
    on the 1M dam break, the 512-scene batch of dam breaks (5.47 M fluid) and a
    5,000-particle dam break.  Host wall clock around the steps, bracketed by stream
    synchronisation."""
    from uph_amd import scenes
    from uph_amd.solver import GpuSolver
    work = {"dam break 1M (999,698 fluid)": lambda: [scenes.dam_break(dx=2 ** 0.5 * 1e-3)],
            "512 dam breaks (5.47 M fluid)": scenes.batch_dam_breaks,
            "dam break 5K": lambda: [scenes.dam_break()]}
    out = {}
    for name, make in work.items():
        scs = make()
        nf = sum(s.n_fluid for s in scs)
        res = {}
        for mode in ("eager", "graph", "persistent", "lazy"):
            g = GpuSolver(scs, use_graph=mode == "graph", sort_every_step=mode != "lazy")
            run = g.step_persistent if mode == "persistent" else g.step
            run(warmup)
            g.sync()
            t0 = time.perf_counter()
            run(steps)
            g.sync()
            ms = (time.perf_counter() - t0) * 1e3 / steps
            res[mode] = {"ms_per_step": ms, "updates_per_s": nf * 1e3 / ms}
            g.close()
            log(f"{name}, {mode}: {ms:.4f} ms per step")
        out[name] = res
    print(json.dumps({"compare_launch": out, "steps": steps, "warmup": warmup}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--settle", type=int, default=200,
                    help="untimed steps before the warm-up steps, so that the clocks and the "
                         "translation caches have settled whatever --warmup is (20 timed steps after 5 "
                         "warm-up steps alone measured 156 us per step against 143 us after 200)")
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of CPU-baseline stepping (0 disables)")
    ap.add_argument("--graph", action="store_true", help="also time hipGraph replay")
    ap.add_argument("--kernel", choices=["cubic", "poly7"], default="cubic",
                    help="smoothing kernel: cubic spline (default) or the reference's PolyOrderSeven")
    ap.add_argument("--scheme", choices=["config", "ref"], default="config",
                    help="config (default): Tait pressure and artificial viscosity; ref: the reference's "
                         "run-phase scheme (pressure transport, viscous forces, fluctuations) with its "
                         "decks' viscosity0 1e-6 and epsilon 1 and thermalDiffusivity0 0.0015 (the decks' "
                         "0.015 is explicitly unstable at this resolution)")
    ap.add_argument("--stress-type", type=int, choices=[1, 2], default=1,
                    help="ref scheme: stressType 1 (single-step viscous forces, the decks' choice) or 2 "
                         "(ClassicShearStress + ViscousForcesTwoStep; sorts every step)")
    ap.add_argument("--h-over-dx", type=float, default=1.3,
                    help="smoothing length in lattice spacings (1.3 default; the reference's decks "
                         "use 2.691, where the 40-neighbour cap binds for every particle)")
    ap.add_argument("--vtype", type=int, default=0,
                    help="virtualType of the tank (mirror particles of derivatives.c:3193-4192; 0: none; "
                         "8: the dam break's left, bottom, corner and right images)")
    ap.add_argument("--sort-every-step", action="store_true",
                    help="re-sort the particles into their cells every step (default: the lazy re-sort, "
                         "which re-sorts a scene only when a particle has left its cell by more than the "
                         "cells' growth)")
    ap.add_argument("--skin", type=float, default=0.0,
                    help="lazy re-sort: cells grow by this fraction (0: the library's default, 0.02 for one "
                         "scene)")
    ap.add_argument("--compare-launch", action="store_true",
                    help="only compare per-kernel launches, hipGraph replay and the persistent kernel "
                         "(compare_launch) and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one D2D copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.compare_launch:
        compare_launch(args.steps, args.warmup)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(scenes would shard across GPUs with no communication)", file=sys.stderr)
        sys.exit(2)
    from uph_amd import scenes
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(width=1.0, height=2.0, dx=2 ** 0.5 * 1e-3)
    sc.kernel = scenes.KERNEL_POLY7 if args.kernel == "poly7" else scenes.KERNEL_CUBIC
    if args.h_over_dx != 1.3:
        sc.h = args.h_over_dx * sc.dx
        sc.dt = 0.25 * sc.h / sc.c0
    if args.scheme == "ref":
        sc.mode = scenes.MODE_REF
        # the pressure diffusion is explicit: with the decks' 0.015 m^2/s its SPH Laplacian
        # needs dt < h^2 / (10 gamma kappa) = 3.2 us at h = 1.84 mm, below this dt of 7.3 us
        sc.kappa, sc.mu, sc.eps, sc.alpha_av = 0.0015, 1e-6, 1.0, 0.0
        sc.stress_type = args.stress_type
    stress2 = args.scheme == "ref" and args.stress_type == 2
    # (for the kernel's name in the output) every query expected to be capped and no
    # mirror rows: the handle's first pair kernel is the K-th key search alone
    kth_all = 4 * math.pi * args.h_over_dx ** 2 > sc.K and not args.vtype and not stress2
    if args.vtype:
        sc.vtype, sc.flat_bottom, sc.depth = args.vtype, 4.0, 2.0   # the tank's width, the column's height
    log(f"workload: {sc.n_fluid} fluid + {sc.n_boundary} wall particles, {args.kernel} kernel, "
        f"h = {args.h_over_dx} dx, {args.scheme} scheme, virtualType {args.vtype}")
    g = GpuSolver(sc, sort_every_step=True if args.sort_every_step else None, skin=args.skin)
    g.step(args.settle + args.warmup)
    g.sync()
    lazy = g.resorts() >= 0
    r0 = g.resorts()
    # the timed region: K steps with no event between the kernels
    t0 = time.perf_counter()
    ms_total = g.time_loop(args.steps)
    overflow = g.sync()
    wall = time.perf_counter() - t0
    resorts_timed = g.resorts() - r0 if lazy else args.steps
    # then K more steps with events between the phases, for the per-kernel split
    ph = g.time_phases(args.steps)
    g.sync()
    ms_pair = ph["pair"]
    ms_step = ms_total / args.steps
    value = sc.n_fluid * args.steps / (ms_total * 1e-3)
    # kernel level: the compulsory bytes of one step over the compacted pair pass alone;
    # step level: the same bytes at the measured step rate
    achieved = BYTES_PER_UPDATE * sc.n_fluid / (ms_pair * 1e-3) / 1e9
    step_achieved = BYTES_PER_UPDATE * value / 1e9
    post = "k_post (capped pass and re-sort when needed)" if lazy else "capped pass"
    log(f"timed {args.steps} steps: {ms_total:.3f} ms (host wall {wall*1e3:.1f} ms), {resorts_timed} re-sorts; "
        f"per step: sort {ph['sort']*1e3:.1f} us, pair {ms_pair*1e3:.1f} us, {post} {ph['exact']*1e3:.1f} us")
    out = {
        "metric": "particle-updates/s, 1M-particle dam break (plain path)",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair terms summed in fp64, fp64 integration",
        "data": "synthetic (generated dam-break lattice)",
        "config": {"workload": "dam break 1.0x2.0 m column, dx=sqrt(2) mm, 4x3 m tank",
                   "fluid_particles": sc.n_fluid, "wall_particles": sc.n_boundary,
                   "kernel": args.kernel, "scheme": args.scheme, "h_over_dx": args.h_over_dx, "dt": sc.dt,
                   "stress_type": args.stress_type if args.scheme == "ref" else None,
                   "virtual_type": args.vtype, "capped_steps": overflow,
                   "settle_steps": args.settle,
                   "resort": (f"lazy, cells grown by {g.skin():g}" if lazy
                              else "every step"),
                   "resorts_in_timed_steps": resorts_timed},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "frac_basis": "kernel level: 80 B x fluid particles per launch / mean duration of "
                                   "the compacted pair pass (HIP events); step_frac: 80 B x "
                                   "particle-updates/s of the whole step",
                     "step_achieved": step_achieved, "step_frac": step_achieved / HBM_PEAK_GBS,
                     # the committed counter run covers the default workload only
                     "traffic": (traffic_from_profiles() if args.kernel == "cubic" and lazy and
                                 args.h_over_dx == 1.3 and args.scheme == "config" and not args.vtype
                                 and args.skin <= 0 else None),
                     # stressType 2 runs the stress pass inside the pair phase, and its pair
                     # kernel is the mode-2 instance; a scene whose every query is expected
                     # to be capped (more than K in range on the lattice, no mirror rows)
                     # runs the K-th key search alone first, and all its pair sums in the
                     # capped pass
                     "kernel": (f"k_kth<true> ({args.kernel}, {args.scheme}, one scene; K-th keys only, "
                                f"the pair sums in the capped pass)" if kth_all else
                                (f"k_stress<{int(sc.kernel)}, true> + " if stress2 else "") +
                                f"k_pair_compact<256, {int(sc.kernel)}, {2 if stress2 else int(sc.mode)}, true, "
                                f"{'true' if 4 * math.pi * args.h_over_dx ** 2 > 0.75 * sc.K else 'false'}, "
                                f"{'true' if lazy else 'false'}> ({args.kernel}, {args.scheme}, one scene"
                                f"{', lazy re-sort' if lazy else ''})"),
                     "kernel_ms": ms_pair, "sort_ms": ph["sort"], "capped_pass_ms": ph["exact"],
                     "capped_pass_kernel": post},
    }
    if args.graph:
        gg = GpuSolver(sc, use_graph=True)
        gg.step(args.warmup)
        gg.sync()
        t0 = time.perf_counter()
        gg.step(args.steps)
        gg.sync()
        out["graph_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / args.steps
        gg.close()
    g.close()
    if args.cpu_budget > 0:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
