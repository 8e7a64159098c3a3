"""GPU parity of the generated UMIST network (hydrogen-helium lines of
RATE12, written by ChemicalNetwork.write_solver into tests/generated/umist):
its species sort after the energy term, so the device code works in its own
order (energy last) behind the caller's rows; evaluation and the
reference-faithful integration (both schedules) are bit for bit the oracle's
mass-action path on the same rows."""
import importlib.util
import os

import numpy as np
import pytest

import oracle
from dengo_amd import runtime
from dengo_amd.ics import tolerances

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "tests", "generated", "umist")
KEYS = ("state", "atol", "status", "nsub", "nrej", "nsweep")


@pytest.fixture(scope="module")
def umist():
    spec = importlib.util.spec_from_file_location("umist_run", os.path.join(GEN, "umist_run.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    s = mod.load()
    from dengo_amd.networks import umist_network
    net = umist_network(os.path.join(ROOT, "tests", "golden", "rate12_hhe.txt"))
    yield mod, s, oracle.ma_net_from_network(net)
    s.close()


def _cells(mod, n, seed):
    """Number densities (log-uniform, molecular and neutral H dominant), de
    from charge neutrality, ge for T log-uniform in 20-3000 K; returns
    (mass-density rows, number-density rows)."""
    rng = np.random.default_rng(seed)
    sp = mod.SPECIES
    w = np.array([mod.WEIGHTS[s] for s in sp])
    num = np.zeros((n, len(sp)))
    for k, s in enumerate(sp):
        if s in ("ge", "de"):
            continue
        lo, hi = (1e-2, 1e2) if s in ("us_H_1", "us_H2_1", "us_He_1") else (1e-8, 1e-3)
        num[:, k] = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    fe = {"us_H_0": -1, "us_H_2": 1, "us_H2_2": 1, "us_H3_2": 1, "us_HeH_2": 1, "us_He_2": 1}
    num[:, sp.index("de")] = sum(num[:, sp.index(k)] * v for k, v in fe.items())
    T = np.exp(rng.uniform(np.log(20.0), np.log(3000.0), n))
    chem = [k for k, s in enumerate(sp) if s not in ("ge", "de")]
    dens = (num[:, chem] * w[chem]).sum(axis=1)
    gsum = sum(num[:, k] for k, s in enumerate(sp) if s != "ge") / (5.0 / 3.0 - 1.0)
    num[:, sp.index("ge")] = T * 1.3806504e-16 * gsum / (dens * 1.67e-24)
    mass = num.copy()
    for k, s in enumerate(sp):
        if s != "ge":
            mass[:, k] = num[:, k] * w[k] * 1.67e-24
    return np.ascontiguousarray(mass), num


def test_umist_eval_bitwise(umist):
    mod, s, m = umist
    print("umist eval", flush=True)
    _, num = _cells(mod, 600, 81)
    g = s.eval(num)
    o = oracle.ma_eval_external(m, num, s.tables, s.T_bounds, flags=oracle.PORTABLE_LOG)
    for k, (a, b) in enumerate(zip(g[:5], o[:5])):
        if not np.array_equal(a, b):
            bad = np.argwhere(np.asarray(a) != np.asarray(b))
            pytest.fail("output %d differs at %d entries, first %s: %r vs %r" % (
                k, len(bad), bad[:3].tolist(), np.asarray(a)[tuple(bad[0])],
                np.asarray(b)[tuple(bad[0])]))


@pytest.mark.parametrize("schedule", [runtime.PER_CELL, runtime.PERSISTENT])
def test_umist_evolve_bitwise(umist, schedule):
    mod, s, m = umist
    print("umist evolve, schedule %d" % schedule, flush=True)
    x, _ = _cells(mod, 300, 82)
    rt, at = tolerances(x)
    ref = oracle.ma_evolve_external(m, x, rt, at, s.tables, 1e12, s.T_bounds,
                                    flags=oracle.PORTABLE_LOG, max_substeps=3000)
    got = s.evolve(x, rt, at, 1e12, schedule=schedule, precision=runtime.FP64,
                   max_substeps=3000)
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k
    assert np.mean(ref["status"] == 0) > 0.5
