"""The autotuner's accuracy measure, cache key and precision rule on the CPU
(torch tensors)."""
import torch

from dengo_amd import autotune
from dengo_amd.autotune import agreement


def test_agreement_counts_status_and_relative_error():
    ref = torch.ones(4, 3, dtype=torch.float64)
    st_ref = torch.tensor([0, 0, 0, 1], dtype=torch.int32)
    x = ref.clone()
    x[1, 2] = 1.0 + 2e-6          # outside a 1e-6 bar
    x[2, 0] = 1.0 + 5e-7          # inside
    st = torch.tensor([0, 0, 0, 2], dtype=torch.int32)
    sa, wb = agreement(ref, st_ref, x, st, 1e-6)
    assert sa == 0.75             # the last cell's status differs
    assert abs(wb - 2.0 / 3.0) < 1e-12   # over the three cells both runs finish
    # nothing finished in both runs: the bar holds vacuously
    assert agreement(ref, st_ref + 1, x, st_ref + 1, 1e-6) == (1.0, 1.0)


def test_cache_digest_covers_tables_and_tolerances():
    from types import SimpleNamespace

    import numpy as np

    from dengo_amd.autotune import _digest
    tab = np.ones((3, 8))
    s1 = SimpleNamespace(tables=tab, T_bounds=(1.0, 1e9))
    s2 = SimpleNamespace(tables=tab * 2, T_bounds=(1.0, 1e9))
    rt = torch.full((4, 3), 1e-9, dtype=torch.float64)
    at = torch.full((4, 3), 1e-20, dtype=torch.float64)
    d = _digest(s1, rt, at, None)
    assert d == _digest(s1, rt.clone(), at.clone(), None)
    assert d != _digest(s2, rt, at, None)                 # other tables
    assert d != _digest(s1, rt * 10, at, None)            # other rtol
    assert d != _digest(s1, rt, at * 10, None)            # other atol scale
    assert d != _digest(SimpleNamespace(tables=tab, T_bounds=(1e4, 1e9)), rt, at, None)


def test_choice_prefers_accuracy_within_the_margin():
    def t(ms, ok=True):
        return {"kernel_ms": ms, "accepted": ok}
    # a near-tie goes to the more accurate sweep
    assert autotune.choose_precision({"fp64": t(3000), "mixed": t(1407),
                                      "fp64-fast": t(1424)}) == "fp64-fast"
    # a clear gain goes to the faster one
    assert autotune.choose_precision({"fp64": t(233), "mixed": t(33),
                                      "fp64-fast": t(82)}) == "mixed"
    # rejected variants never win, however fast
    assert autotune.choose_precision({"fp64": t(100), "mixed": t(10, False),
                                      "fp64-fast": t(99)}) == "fp64"
    assert autotune.choose_precision({"fp64": t(100), "fp64-fast": t(96)}) == "fp64"
    assert autotune.choose_precision({"fp64": t(100), "fp64-fast": t(96)}, margin=0.0) == \
        "fp64-fast"


def test_schedule_candidates():
    from dengo_amd import runtime
    c = dict(autotune.schedule_candidates(256, dense=False))
    assert c["persistent"] == dict(schedule=runtime.PERSISTENT, grid=0)
    assert [c["persistent_%d_per_cu" % k]["grid"] for k in (4, 2, 1)] == [1024, 512, 256]
    assert c["per_cell"] == dict(schedule=runtime.PER_CELL)
    d = dict(autotune.schedule_candidates(256, dense=True))
    assert sorted(d) == ["per_cell", "persistent", "persistent_0.5_per_cu"]
    assert d["persistent_0.5_per_cu"]["grid"] == 128


def test_agreement_ignores_trace_species_but_not_the_energy():
    ref = torch.tensor([[1.0, 1e-25, 5.0]], dtype=torch.float64)   # species, trace species, energy
    st = torch.zeros(1, dtype=torch.int32)
    species = torch.tensor([True, True, False])
    x = ref.clone()
    x[0, 1] = 3e-25                  # a species at the floor: relative error 2, not counted
    assert agreement(ref, st, x, st, 1e-6, species) == (1.0, 1.0)
    assert agreement(ref, st, x, st, 1e-6) == (1.0, 1.0)    # every column a species
    y = x.clone()
    y[0, 0] = 1.0 + 2e-6             # a major species is counted
    assert agreement(ref, st, y, st, 1e-6, species) == (1.0, 0.0)
    x[0, 2] = 5.0 * (1 + 1e-5)       # the energy is always counted
    assert agreement(ref, st, x, st, 1e-6, species) == (1.0, 0.0)


def test_species_mask():
    from types import SimpleNamespace
    m = autotune.species_mask(SimpleNamespace(N=4))
    assert m.tolist() == [True, True, True, False]
    m = autotune.species_mask(SimpleNamespace(N=4, energy_column=1))
    assert m.tolist() == [True, False, True, True]
