"""Autotuner for the launch and cell parameters of the stepping path.

For a given workload (one scene or a batch) it times every combination of

  pair_variant  0 (two-pass compacted kernel) / 1 (direct loop)
  pair_block    work-group size of the compacted kernel: 128, 256, 512
  cell_factor   cell edge as a multiple of the search radius: 1.0, 1.15, 1.3

for a few steps each, with HIP events on the library's stream, and returns the
fastest.  Larger cells hold more candidates per row but fewer cells to scan;
the best trade-off depends on particle count and on how the scenes of a batch
are sized.  Results can be cached in a JSON file keyed by the workload's shape.
"""
from __future__ import annotations

import itertools
import json
import os
from typing import Callable, Dict, List, Optional, Sequence

from . import scenes as _scenes
from .scenes import Scene

DEFAULT_SPACE = {
    "pair_variant": (0, 1),
    "pair_block": (128, 256, 512),
    "cell_factor": (1.0, 1.15, 1.3),
}


def workload_key(scenes: Sequence[Scene]) -> str:
    nf = sum(s.n_fluid for s in scenes)
    nb = sum(s.n_boundary for s in scenes)
    hs = sorted({round(s.h / s.dx, 4) for s in scenes})
    return f"scenes={len(scenes)},nf={nf},nb={nb},h/dx={hs}"


def candidates(space: Optional[Dict] = None) -> List[Dict]:
    space = space or DEFAULT_SPACE
    out = []
    for var, blk, cf in itertools.product(space["pair_variant"], space["pair_block"],
                                          space["cell_factor"]):
        if var == 1 and blk != space["pair_block"][0]:
            continue          # the direct kernel has a fixed work-group size
        out.append({"pair_variant": var, "pair_block": blk, "cell_factor": cf})
    return out


def pick_best(timings: List[Dict]) -> Dict:
    """The fastest configuration among those whose runs were valid."""
    ok = [t for t in timings if t.get("ms_per_step") is not None and t.get("error") is None]
    if not ok:
        raise RuntimeError("autotune: no configuration ran")
    return min(ok, key=lambda t: t["ms_per_step"])


def _time_config(scenes, cfg, warmup, steps) -> float:
    from .solver import GpuSolver
    g = GpuSolver(scenes, **cfg)
    try:
        g.step(warmup)
        g.sync()
        tot, _ = g.time_steps(steps)
        return tot / steps
    finally:
        g.close()


def autotune(scenes: Sequence[Scene] | Scene, warmup: int = 3, steps: int = 10,
             space: Optional[Dict] = None, cache: Optional[str] = None,
             timer: Optional[Callable] = None, log: Optional[Callable[[str], None]] = None) -> Dict:
    """Time each candidate configuration; return {'best': cfg, 'timings': [...]}."""
    if isinstance(scenes, Scene):
        scenes = [scenes]
    key = workload_key(scenes)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            db = json.load(f)
        if key in db:
            return db[key]
    timer = timer or _time_config
    timings = []
    for cfg in candidates(space):
        rec = dict(cfg)
        try:
            rec["ms_per_step"] = float(timer(scenes, cfg, warmup, steps))
            rec["error"] = None
        except RuntimeError as e:          # a configuration the library rejects
            rec["ms_per_step"] = None
            rec["error"] = str(e)
        timings.append(rec)
        if log:
            log(f"autotune {cfg}: {rec['ms_per_step']}")
    best = pick_best(timings)
    res = {"best": {k: best[k] for k in ("pair_variant", "pair_block", "cell_factor")},
           "ms_per_step": best["ms_per_step"], "timings": timings, "workload": key}
    if cache:
        db = {}
        if os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
        db[key] = res
        with open(cache, "w") as f:
            json.dump(db, f, indent=1)
    return res


WORKLOADS = {
    "dam1m": lambda: _scenes.dam_break(dx=2 ** 0.5 * 1e-3),          # 999,698 fluid
    "batch512": lambda: _scenes.batch_dam_breaks(),                   # 512 scenes, 5.47 M fluid
    "dam5k": lambda: _scenes.dam_break(),
    "wave240k": lambda: _scenes.wave_tank(),                          # 237,900 fluid, piston
}


def main(argv=None):
    """Tune one named workload, then time the pick against the library's defaults over
    the same steps (one launch per kernel, device-synchronised wall clock)."""
    import argparse
    import sys
    import time
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="batch512")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cache", default=None)
    ap.add_argument("--no-tune", action="store_true",
                    help="only time the library's defaults (for a kernel trace of the workload)")
    ap.add_argument("--sort-every-step", action="store_true",
                    help="time handles that re-sort every step (default: the library's choice, lazy "
                         "re-sort for at most 16 scenes)")
    ap.add_argument("--lazy", action="store_true",
                    help="time handles that re-sort lazily whatever the batch size")
    ap.add_argument("--skin", type=float, default=0.0,
                    help="lazy re-sort: cells grow by this fraction (0: the library's default)")
    a = ap.parse_args(argv)
    from .solver import GpuSolver
    scs = WORKLOADS[a.workload]()
    scs = [scs] if isinstance(scs, Scene) else scs
    nf = sum(s.n_fluid for s in scs)

    def say(m):
        print(m, file=sys.stderr, flush=True)

    say(f"{a.workload}: {len(scs)} scene(s), {nf} fluid")
    out = {"workload": a.workload, "n_fluid": nf}
    legs = [("default", {})]
    if not a.no_tune:
        res = autotune(scs, warmup=a.warmup, steps=a.steps, cache=a.cache, log=say)
        out.update(best=res["best"], tuned_ms=res["ms_per_step"])
        legs.append(("best", res["best"]))
    for name, cfg in legs:
        ses = True if a.sort_every_step else (False if a.lazy else None)
        g = GpuSolver(scs, sort_every_step=ses, skin=a.skin, **cfg)
        g.step(a.warmup)
        g.sync()
        r0 = g.resorts()
        t0 = time.perf_counter()
        g.step(100)
        g.sync()
        ms = (time.perf_counter() - t0) * 10.0
        nres = g.resorts() - r0 if r0 >= 0 else -1
        g.close()
        out[name + "_ms_per_step"] = ms
        out[name + "_updates_per_s"] = nf * 1e3 / ms
        out[name + "_resorts_in_100_steps"] = nres
        say(f"{name} {cfg}: {ms:.4f} ms per step, {nf * 1e3 / ms:.3e} updates/s, "
            f"{'sorted every step' if nres < 0 else f'{nres} re-sorts'}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
