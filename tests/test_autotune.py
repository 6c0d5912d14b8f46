"""Autotuner: candidate space and selection (CPU, with a stand-in timer)."""
import json

import pytest

from uph_amd import autotune, scenes


def test_candidates_cover_space():
    c = autotune.candidates()
    assert {"pair_variant": 0, "pair_block": 512, "cell_factor": 1.3} in c
    assert all(x["pair_block"] == 128 for x in c if x["pair_variant"] == 1)
    assert len(c) == 3 * 3 + 3


def test_pick_best_skips_failures():
    t = [{"ms_per_step": 0.3, "error": None, "pair_variant": 0, "pair_block": 256, "cell_factor": 1.0},
         {"ms_per_step": None, "error": "x", "pair_variant": 1, "pair_block": 128, "cell_factor": 1.0},
         {"ms_per_step": 0.2, "error": None, "pair_variant": 0, "pair_block": 128, "cell_factor": 1.3}]
    assert autotune.pick_best(t)["pair_block"] == 128
    with pytest.raises(RuntimeError):
        autotune.pick_best([t[1]])


def test_autotune_with_timer_and_cache(tmp_path):
    sc = scenes.dam_break(width=0.2, height=0.2)
    cost = lambda s, cfg, w, n: 1.0 + cfg["pair_variant"] + abs(cfg["cell_factor"] - 1.15) \
        + (0.1 if cfg["pair_block"] != 512 else 0.0)
    cache = tmp_path / "tune.json"
    r = autotune.autotune(sc, timer=cost, cache=str(cache))
    assert r["best"] == {"pair_variant": 0, "pair_block": 512, "cell_factor": 1.15}
    db = json.loads(cache.read_text())
    assert autotune.workload_key([sc]) in db
    # cached: the timer is not called again
    r2 = autotune.autotune(sc, timer=lambda *a: 1 / 0, cache=str(cache))
    assert r2["best"] == r["best"]
