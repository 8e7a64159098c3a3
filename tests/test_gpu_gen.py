"""GPU parity of generated device networks (gen_net.h; RHS and Jacobian
printed from the sympy forms by dengo_amd.codegen):

* primgen, the full primordial network through the generator, primphoto,
  the primordial species with a reduced reaction and cooling set plus
  photoionisation and photoheating on redshift tables, and oxydata, O I-IX
  with the reference's Gnat & Ferland cooling and HM12 photoionisation and
  photoheating tables, evaluate (T, rates,
  RHS, Jacobian) and integrate (reference-faithful FP64, both schedules) bit
  for bit like the oracle's generated-network path, with the redshift inside
  and outside the tables' range;
* the fast FP64 and mixed sweeps end in the same status as the faithful one
  on almost every cell;
* the generated run_<name> harness runs the reference's Python loop.
The oracle's path and the generated expressions are pinned against the
hand-written primordial restatement and numpy on the CPU
(tests/test_codegen_gen.py)."""
import importlib.util
import os

import numpy as np
import pytest

import oracle
from dengo_amd import runtime
from dengo_amd.ics import oxygen_ics, primordial_ics, tolerances

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "tests", "generated")
KEYS = ("state", "atol", "status", "nsub", "nrej", "nsweep")


def _module(name):
    spec = importlib.util.spec_from_file_location(
        "%s_run" % name, os.path.join(GEN, name, "%s_run.py" % name))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _cells(name, n, seed, **kw):
    """(mass-density input, number densities) of cells for network `name`."""
    if name == "oxydata":
        x, num, _ = oxygen_ics(n, seed=seed, T_range=kw.get("T_range", (1e3, 1e8)))
    else:
        x, num, _ = primordial_ics(n, seed=seed, **kw)
    return x, num


@pytest.fixture(scope="module", params=["primgen", "primphoto", "oxydata"])
def gen(request):
    mod = _module(request.param)
    s = mod.load()
    yield request.param, mod, s
    s.close()


@pytest.mark.parametrize("z", [-1.0, 2.5, 9.99, 10.0])
def test_generated_eval_bitwise(gen, z):
    name, _, s = gen
    print("generated eval: %s z=%g" % (name, z), flush=True)
    _, n = _cells(name, 700, 61, T_range=(3.0, 8e8))
    g = s.eval(n, z=z)
    o = oracle.gen_eval(oracle.gen_net(name), n, s.tables, s.T_bounds, z=z,
                        flags=oracle.PORTABLE_LOG)
    for k, (a, b) in enumerate(zip(g[:5], o[:5])):
        assert np.array_equal(a, b), (name, z, k)


@pytest.mark.parametrize("schedule", [runtime.PER_CELL, runtime.PERSISTENT])
def test_generated_evolve_bitwise(gen, schedule):
    name, _, s = gen
    print("generated evolve: %s schedule %d" % (name, schedule), flush=True)
    x, _ = _cells(name, 240, 62)
    rt, at = tolerances(x)
    z = -1.0 if name == "primgen" else 3.0
    ref = oracle.gen_evolve(oracle.gen_net(name), x, rt, at, s.tables, 1e12, s.T_bounds, z=z,
                            flags=oracle.PORTABLE_LOG, max_substeps=3000)
    got = s.evolve(x, rt, at, 1e12, z=z, schedule=schedule, precision=runtime.FP64,
                   max_substeps=3000)
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), (name, k)
    assert got["ttot"] == ref["ttot"] and got["dt"] == ref["dt"]


@pytest.mark.parametrize("precision,form", [(runtime.FP64_FAST, 0), (runtime.FP64_FAST, 1),
                                            (runtime.MIXED, 0)])
def test_generated_fast_sweeps_agree_with_faithful(gen, precision, form):
    """form 0 runs the FP64-fast sweep as the register-lean kernel where the
    network's compiled and fits (fused reaction loop), form 1 as the
    register-resident one."""
    name, _, s = gen
    print("generated fast sweeps: %s precision %d form %d" % (name, precision, form), flush=True)
    x, _ = _cells(name, 2000, 63)
    rt, at = tolerances(x)
    z = -1.0 if name == "primgen" else 3.0
    ref = s.evolve(x, rt, at, 1e11, z=z, precision=runtime.FP64, max_substeps=3000)
    got = s.evolve(x, rt, at, 1e11, z=z, precision=precision, max_substeps=3000,
                   kernel_form=form)
    assert np.mean(got["status"] == ref["status"]) >= 0.97
    done = (got["status"] == 0) & (ref["status"] == 0)
    rel = np.abs(got["state"] - ref["state"]) / np.maximum(np.abs(ref["state"]), 1e-300)
    nsp = x.shape[1] - 2
    big = ref["state"] > 1e-10 * ref["state"][:, :nsp].sum(axis=1, keepdims=True)
    err = np.where(big, rel, 0.0)[done].max(axis=1)
    assert np.mean(err <= 1e-6) >= 0.97


def test_generated_run_harness(gen):
    name, mod, s = gen
    print("generated run harness: %s" % name, flush=True)
    _, n = _cells(name, 6, 64)
    species = mod.SPECIES
    ics = {sp: n[:, k] * mod.WEIGHTS[sp] for k, sp in enumerate(species)}
    rv, rv_t = getattr(mod, "run_%s" % name)(ics, 1e9, niter=200, solver=s)
    assert set(species) <= set(rv)
    ok = rv_t["successful"]
    assert ok.any()
    t = rv_t["t"][ok]
    assert np.all(np.diff(t) > 0) and t[-1] < 1e9
    assert all(np.all(np.isfinite(rv[sp])) for sp in species)
