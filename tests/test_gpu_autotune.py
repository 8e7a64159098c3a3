"""The autotuner on the GPU: on samples of the primordial and oxygen
workloads every fast sweep meets the accuracy bar (the mixed sweep solves the
primordial network's cells with species at the 1e-25 floor with the FP64 LU,
so it keeps the reference's capped cells); the choice is the most accurate
accepted variant within 5% of the fastest; a cached result is returned
without re-running."""
import json
import os

import pytest
import torch

from dengo_amd import autotune, runtime
from dengo_amd.ics import oxygen_ics, primordial_ics, tolerances
from dengo_amd.networks import oxygen_tables

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _dev(*arrays):
    return [torch.from_numpy(a).to(DEV) for a in arrays]


def test_primordial_and_oxygen_choices(solver, tmp_path):
    cache = str(tmp_path / "tune.json")
    inp, _, _ = primordial_ics(4096, seed=2024)
    x, rt, at = _dev(inp, *tolerances(inp))
    print("autotune primordial, 2048-cell sample", flush=True)
    r = autotune.tune(solver, x, rt, at, 1e12, sample=2048, cache=cache)
    trials = r["precision_trials"]
    for v in ("mixed", "fp64-fast"):
        assert trials[v]["accepted"] and trials[v]["status_agreement"] >= 0.99, trials
    assert r["precision"] == autotune.choose_precision(trials)
    assert r["options"]["precision"] == {"fp64": runtime.FP64, "mixed": runtime.MIXED,
                                         "fp64-fast": runtime.FP64_FAST}[r["precision"]]
    assert r["schedule"] in r["schedule_ms"]
    # the options run as given
    y = x[:64].clone()
    solver.evolve_device(y, rt[:64].contiguous(), at[:64].clone(), 1e12, **r["options"])
    # cached: same answer from the file
    saved = json.load(open(cache))
    assert r["key"] in saved
    assert autotune.tune(solver, x, rt, at, 1e12, sample=2048, cache=cache) == saved[r["key"]]

    oxy = runtime.OxygenSolver(oxygen_tables()[1])
    try:
        xo, _, _ = oxygen_ics(4096, seed=2025)
        x2, rt2, at2 = _dev(xo, *tolerances(xo))
        print("autotune oxygen, 4096-cell sample", flush=True)
        r2 = autotune.tune(oxy, x2, rt2, at2, 1e12, sample=4096, cache=cache)
        trials = r2["precision_trials"]
        assert trials["mixed"]["accepted"] and trials["mixed"]["within_bar"] >= 0.99, trials
        assert r2["precision"] == autotune.choose_precision(trials)
    finally:
        oxy.close()
    assert len(json.load(open(cache))) == 2
