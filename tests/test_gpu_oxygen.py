"""GPU parity of the oxygen ion-by-ion network against the CPU oracle.

The device network is generated by dengo_amd.codegen (net_oxygen.h) and run
by the same kernels as the primordial network; the reference-faithful FP64
kernel must match the oracle's portable-log mode bit for bit."""
import os

import numpy as np
import pytest

import oracle
from dengo_amd import runtime
from dengo_amd.ics import oxygen_ics, tolerances
from dengo_amd.networks import oxygen_tables

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oxygen_cells.npz")
NET = oracle.ion_chain_net(8, 15.9994)


@pytest.fixture(scope="module")
def oxy():
    _, tab = oxygen_tables()
    s = runtime.OxygenSolver(tab)
    yield s, tab
    s.close()


def test_oxygen_eval_bitwise(oxy):
    s, tab = oxy
    _, n, _ = oxygen_ics(1024, seed=13)
    g = s.eval(n)
    o = oracle.ma_eval(NET, n, tab, (1e4, 1e9), flags=oracle.PORTABLE_LOG)
    for a, b in zip(g[:4], o[:4]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("schedule", [runtime.PER_CELL, runtime.PERSISTENT])
def test_oxygen_fp64_bitwise(oxy, schedule):
    s, _ = oxy
    g = np.load(GOLD)
    r = s.evolve(g["input"], g["rtol"], g["atol"], float(g["dtf"]), schedule=schedule)
    for k in ("state", "atol", "status", "nsub", "nrej", "nsweep", "T"):
        assert np.array_equal(r[k], g["plog_" + k]), k
    assert r["ttot"] == float(g["plog_ttot"])


def test_oxygen_fp64_fast(oxy):
    """Structure-exploiting FP64: same status everywhere; on cells that take
    the oracle's substep path (>= 95%) every component with mass fraction above
    1e-10 of its cell agrees to 1e-6, and every component stays within the
    solver's own tolerance (|d| <= atol + rtol |x|)."""
    s, _ = oxy
    g = np.load(GOLD)
    r = s.evolve(g["input"], g["rtol"], g["atol"], float(g["dtf"]),
                 precision=runtime.FP64_FAST)
    ref = g["libm_state"]
    assert np.array_equal(r["status"], g["libm_status"])
    same = (r["nsub"] == g["libm_nsub"]) & (r["nrej"] == g["libm_nrej"])
    assert same.mean() >= 0.95
    frac = ref[:, :9] / ref[:, :9].sum(axis=1, keepdims=True)
    rel = np.abs(r["state"] - ref) / np.abs(ref)
    sig = np.concatenate([frac > 1e-10, np.ones((len(ref), 2), bool)], axis=1)
    assert rel[same][sig[same]].max() < 1e-6
    tolu = np.abs(r["state"] - ref) / (g["atol"] + g["rtol"] * np.abs(ref))
    assert tolu[same].max() <= 1.0


def test_oxygen_mixed_status(oxy):
    """FP32 LU with FP64 iterative refinement and an FP64 fallback (config 3's
    mixed form): same status everywhere; on cells that take the oracle's
    substep path (>= 95%) every component with mass fraction above 1e-10
    agrees to 1e-6, and every component stays within the solver's own
    tolerance."""
    s, _ = oxy
    g = np.load(GOLD)
    r = s.evolve(g["input"], g["rtol"], g["atol"], float(g["dtf"]), precision=runtime.MIXED)
    ref = g["libm_state"]
    assert np.array_equal(r["status"], g["libm_status"])
    assert np.all(np.isfinite(r["state"]))
    same = (r["nsub"] == g["libm_nsub"]) & (r["nrej"] == g["libm_nrej"])
    assert same.mean() >= 0.95
    frac = ref[:, :9] / ref[:, :9].sum(axis=1, keepdims=True)
    rel = np.abs(r["state"] - ref) / np.abs(ref)
    sig = np.concatenate([frac > 1e-10, np.ones((len(ref), 2), bool)], axis=1)
    assert rel[same][sig[same]].max() < 1e-6
    tolu = np.abs(r["state"] - ref) / (g["atol"] + g["rtol"] * np.abs(ref))
    assert tolu[same].max() <= 1.0
