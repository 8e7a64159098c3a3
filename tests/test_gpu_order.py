"""The hand-out order of cells (dengo_options.order): results per cell do not
depend on it, bit for bit, on every network and both schedules, eager and
captured; an order that is not a permutation of the cells is refused before
anything runs; runtime.cost_order builds the costly-first order from a
previous call's sweep counts, runtime.density_order the densest-first order
from the inputs."""
import os

import numpy as np
import pytest
import torch

from dengo_amd import runtime
from dengo_amd.ics import oxygen_ics, primordial_ics, tolerances
from dengo_amd.networks import oxygen_tables
from dengo_amd.synthetic import SyntheticNetwork

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
KEYS = ("state", "atol", "status", "nsub", "nrej", "nsweep")
GOLD_SYN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "synthetic_cells.npz")


@pytest.fixture(scope="module")
def oxy():
    s = runtime.OxygenSolver(oxygen_tables()[1])
    yield s
    s.close()


@pytest.fixture(scope="module")
def dense():
    s = runtime.SyntheticSolver(SyntheticNetwork(seed=7))
    yield s
    s.close()


def _same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("schedule", [runtime.PERSISTENT, runtime.PER_CELL])
@pytest.mark.parametrize("precision", [runtime.FP64, runtime.FP64_FAST])
def test_primordial_order_does_not_change_results(solver, schedule, precision):
    inp, _, _ = primordial_ics(300, seed=41)
    rtol, atol = tolerances(inp)
    print("order test: primordial 300 cells, schedule %d precision %d" % (schedule, precision),
          flush=True)
    ref = solver.evolve(inp, rtol, atol, 1e10, schedule=schedule, precision=precision,
                        max_substeps=4000)
    perm = torch.from_numpy(np.random.default_rng(3).permutation(300)).to(DEV)
    got = solver.evolve(inp, rtol, atol, 1e10, schedule=schedule, precision=precision,
                        max_substeps=4000, order=perm)
    _same(got, ref)
    costly = runtime.cost_order(torch.from_numpy(ref["nsweep"]).to(DEV))
    assert costly.dtype == torch.int64 and costly.numel() == 300
    sw = ref["nsweep"][costly.cpu().numpy()]
    assert np.all(sw[:-1] >= sw[1:])
    _same(solver.evolve(inp, rtol, atol, 1e10, schedule=schedule, precision=precision,
                        max_substeps=4000, order=costly), ref)


def test_orders_that_are_not_permutations_are_refused(solver):
    inp, _, _ = primordial_ics(128, seed=42)
    rtol, atol = tolerances(inp)
    perm = np.random.default_rng(4).permutation(128)
    bad = [perm.copy(), perm.copy(), perm.copy()]
    bad[0][7] = -1              # outside
    bad[1][9] = 128             # outside
    bad[2][3] = bad[2][4]       # repeated
    for b in bad:
        with pytest.raises(RuntimeError, match="not a permutation"):
            solver.evolve(inp, rtol, atol, 1e9, order=torch.from_numpy(b).to(DEV))
    with pytest.raises(ValueError):
        solver.evolve(inp, rtol, atol, 1e9, order=torch.zeros(10, dtype=torch.int64, device=DEV))
    # the solver is still usable, and a valid order gives the natural result
    ref = solver.evolve(inp, rtol, atol, 1e9, max_substeps=4000)
    _same(solver.evolve(inp, rtol, atol, 1e9, max_substeps=4000,
                        order=torch.from_numpy(perm).to(DEV)), ref)


def test_captured_launch_with_order(solver):
    inp, _, _ = primordial_ics(200, seed=44)
    rtol, atol = tolerances(inp)
    ref = solver.evolve(inp, rtol, atol, 1e9, precision=runtime.FP64_FAST, max_substeps=4000)
    order = runtime.cost_order(torch.from_numpy(ref["nsweep"]).to(DEV))
    x = torch.from_numpy(inp).to(DEV)
    rt = torch.from_numpy(rtol).to(DEV)
    at = torch.from_numpy(atol).to(DEV)
    g = runtime.CapturedEvolve(solver, x, rt, at, 1e9, precision=runtime.FP64_FAST,
                               max_substeps=4000, order=order)
    for _ in range(2):
        x.copy_(torch.from_numpy(inp))
        at.copy_(torch.from_numpy(atol))
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(x.cpu().numpy(), ref["state"])
        assert np.array_equal(at.cpu().numpy(), ref["atol"])
    g.close()


def test_two_ordered_graphs_survive_larger_eager_calls(solver):
    # each graph gathers and scatters through buffers of its own: an eager
    # ordered call on more cells (which regrows the solver's buffers) and a
    # second ordered graph between replays leave every replay bitwise equal
    # to the eager result
    sizes = (96, 160)
    graphs, refs, bufs = [], [], []
    for k, n in enumerate(sizes):
        inp, _, _ = primordial_ics(n, seed=60 + k)
        rtol, atol = tolerances(inp)
        ref = solver.evolve(inp, rtol, atol, 1e9, precision=runtime.FP64_FAST, max_substeps=3000)
        order = torch.from_numpy(np.random.default_rng(k).permutation(n)).to(DEV)
        x = torch.from_numpy(inp).to(DEV)
        rt = torch.from_numpy(rtol).to(DEV)
        at = torch.from_numpy(atol).to(DEV)
        graphs.append(runtime.CapturedEvolve(solver, x, rt, at, 1e9, precision=runtime.FP64_FAST,
                                             max_substeps=3000, order=order))
        refs.append(ref)
        bufs.append((inp, atol, x, at))
    big, _, _ = primordial_ics(700, seed=69)
    brt, bat = tolerances(big)
    bref = solver.evolve(big, brt, bat, 1e9, precision=runtime.FP64_FAST, max_substeps=3000)
    border = torch.from_numpy(np.random.default_rng(9).permutation(700)).to(DEV)
    for rep in range(3):
        for g, ref, (inp, atol, x, at) in zip(graphs, refs, bufs):
            x.copy_(torch.from_numpy(inp))
            at.copy_(torch.from_numpy(atol))
            g.launch()
            torch.cuda.synchronize()
            assert np.array_equal(x.cpu().numpy(), ref["state"]), rep
            assert np.array_equal(at.cpu().numpy(), ref["atol"]), rep
        # an eager ordered call on more cells between replays
        _same(solver.evolve(big, brt, bat, 1e9, precision=runtime.FP64_FAST, max_substeps=3000,
                            order=border), bref)
    for g in graphs:
        g.close()


def test_oxygen_order(oxy):
    inp, _, _ = oxygen_ics(256, seed=43)
    rtol, atol = tolerances(inp)
    ref = oxy.evolve(inp, rtol, atol, 1e11, precision=runtime.MIXED)
    perm = torch.from_numpy(np.random.default_rng(5).permutation(256)).to(DEV)
    _same(oxy.evolve(inp, rtol, atol, 1e11, precision=runtime.MIXED, order=perm), ref)
    dense_first = runtime.density_order(torch.from_numpy(inp).to(DEV), 10)
    _same(oxy.evolve(inp, rtol, atol, 1e11, precision=runtime.MIXED, order=dense_first), ref)


def test_dense_order(dense):
    s = dense
    g = np.load(GOLD_SYN)
    sl = slice(0, 6)
    x, T, rt, at = (g[k][sl].copy() for k in ("input", "T", "rtol", "atol"))
    dtf = float(g["dtf"])
    ref = s.evolve(x, T, rt, at, dtf, precision=runtime.FP64_FAST)
    order = torch.tensor([5, 3, 2, 1, 0, 4], dtype=torch.int64, device=DEV)
    print("order test: dense 6 cells", flush=True)
    _same(s.evolve(x, T, rt, at, dtf, precision=runtime.FP64_FAST, order=order), ref)
    order[2] = 6
    with pytest.raises(RuntimeError, match="not a permutation"):
        s.evolve(x, T, rt, at, dtf, precision=runtime.FP64_FAST, order=order)
