"""The launch-shape autotuner: the search on a synthetic objective (CPU) and
one tuning run of a small discharge on the GPU."""
import json

import pytest

from cudapic_amd import autotune


def test_coordinate_descent_finds_separable_optimum():
    space = {"a": [1, 2, 3, 4], "b": [10, 20, 30], "c": [0, 1]}
    best_of = {"a": 3, "b": 30, "c": 1}
    calls = []

    def f(cfg):
        calls.append(cfg)
        return 100.0 - sum(abs(cfg[k] - best_of[k]) for k in cfg)

    best, val, evals = autotune.coordinate_descent(f, space)
    assert best == best_of and val == 100.0
    # each distinct configuration is measured once
    assert len(calls) == len({tuple(sorted(c.items())) for c in calls}) == len(evals)


def test_coordinate_descent_ignores_gains_below_threshold():
    space = {"a": [0, 1], "b": [0, 1]}

    def f(cfg):
        return 1.0 + 0.001 * cfg["a"] + 0.5 * cfg["b"]

    best, val, _ = autotune.coordinate_descent(f, space, min_gain=0.01)
    assert best == {"a": 0, "b": 1}


def test_coordinate_descent_second_pass_follows_interaction():
    # b only pays once a has moved: the second pass must pick it up
    space = {"b": [0, 1], "a": [0, 1]}

    def f(cfg):
        return 1.0 + cfg["a"] + 2.0 * cfg["a"] * cfg["b"]

    best, val, _ = autotune.coordinate_descent(f, space, passes=2)
    assert best == {"a": 1, "b": 1} and val == 4.0


@pytest.mark.gpu
def test_tune_sim_small_discharge(tmp_path, monkeypatch):
    import cudapic_amd as ca
    from helpers import synthetic_state
    monkeypatch.setenv("CUDAPIC_AUTOTUNE_CACHE", str(tmp_path / "cache.json"))
    monkeypatch.setattr(autotune, "SIM_SPACE", {"block_size": [256, 512], "lanes_per_cu": [1024, 512],
                                                "ppt_inplace": [2, 1], "lds_replicas": [0], "replicas": [8],
                                                "segments": [0], "ppt": [4]})
    p = ca.params(weight=2.0e3, precision=ca.PIC_PREC_FP64)
    E, I = synthetic_state(200_000, 210_000, seed=5)
    res = autotune.tune_sim(p, 400_000, E, I, steps=80, warmup=80, passes=1)
    assert res["rate"] > 1e9 and res["evaluations"] == 4
    cfg = res["config"]
    assert cfg["ppt_inplace"] in (1, 2) and cfg["grid_blocks"] > 0 and cfg["block_size"] in (256, 512)
    # cached: the second call measures nothing
    again = autotune.tune_sim(p, 400_000, E, I)
    assert again == json.loads(json.dumps(res))
    # the tuned configuration runs
    with ca.Sim(p, capacity=400_000, **cfg) as sim:
        sim.upload(0, *E)
        sim.upload(1, *I)
        sim.run(80)
        sim.check()
        assert sim.counts()[0] > 150_000
