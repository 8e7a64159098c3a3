"""GPU parity of the solver template set and the strip solve.

* A network this package has no hand-written code for (carbon, C I-C VII,
  written by ChemicalNetwork.write_solver into tests/generated/carbon and
  built by dengo_amd.build.build_solver) integrates bit for bit like the
  oracle's runtime-described mass-action network.
* dengo_<n>_be_chem_solve (BE_chem_solve on a strip of cells sharing dt and
  the convergence decision) matches the oracle's strip solve bit for bit,
  and run_<n> (the Cython harness's loop) matches the same loop over the
  oracle."""
import importlib.util
import os

import numpy as np
import pytest

import oracle
from dengo_amd import harness, runtime
from dengo_amd.ics import ion_ics, primordial_ics, tolerances
from dengo_amd.networks import PRIMORDIAL_SPECIES

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CARBON = os.path.join(ROOT, "tests", "generated", "carbon")
C_WEIGHT = 12.0107
C_NET = oracle.ion_chain_net(6, C_WEIGHT)
P_WEIGHTS = dict(zip(PRIMORDIAL_SPECIES, [2.0, 2.0, 1.00794, 1.00794, 1.00794, 4.002602,
                                          4.002602, 4.002602, 1.0, 1.0]))


def carbon_module():
    spec = importlib.util.spec_from_file_location("carbon_run",
                                                  os.path.join(CARBON, "carbon_run.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def carbon():
    mod = carbon_module()
    s = mod.load()
    yield mod, s
    s.close()


def test_carbon_eval_bitwise(carbon):
    _, s = carbon
    _, n, _ = ion_ics(512, 41, 7, C_WEIGHT)
    g = s.eval(n)
    o = oracle.ma_eval(C_NET, n, s.tables, s.T_bounds, flags=oracle.PORTABLE_LOG)
    for a, b in zip(g[:4], o[:4]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("schedule", [runtime.PER_CELL, runtime.PERSISTENT])
def test_carbon_evolve_bitwise(carbon, schedule):
    _, s = carbon
    x, _, _ = ion_ics(300, 42, 7, C_WEIGHT)
    rt, at = tolerances(x)
    r = s.evolve(x, rt, at, 1e12, schedule=schedule)
    o = oracle.ma_evolve(C_NET, x, rt, at, s.tables, 1e12, s.T_bounds, flags=oracle.PORTABLE_LOG)
    for k in ("state", "atol", "status", "nsub", "nrej", "nsweep", "T"):
        assert np.array_equal(r[k], o[k]), k
    assert r["ttot"] == o["ttot"]


@pytest.mark.parametrize("form", [0, 1])
def test_carbon_fp64_fast_agrees_with_faithful(carbon, form):
    """The FP64-fast sweep of a descriptor network: form 0 runs the
    register-lean kernel (its reactions fused into one loop, so sums round in
    another order), form 1 the register-resident one.  Against the faithful
    sweep (bit for bit the oracle's, above): the same status on 97% of the
    cells and, on 97% of the cells both finish, the energy and every species
    above 1e-10 of the cell's total within 1e-6."""
    _, s = carbon
    x, _, _ = ion_ics(2000, 43, 7, C_WEIGHT)
    rt, at = tolerances(x)
    print("carbon fp64fast form %d" % form, flush=True)
    ref = s.evolve(x, rt, at, 1e11, precision=runtime.FP64, max_substeps=3000)
    got = s.evolve(x, rt, at, 1e11, precision=runtime.FP64_FAST, max_substeps=3000,
                   kernel_form=form)
    assert np.mean(got["status"] == ref["status"]) >= 0.97
    done = (got["status"] == 0) & (ref["status"] == 0)
    assert done.mean() > 0.5
    rel = np.abs(got["state"] - ref["state"]) / np.maximum(np.abs(ref["state"]), 1e-300)
    nsp = x.shape[1] - 1
    big = ref["state"] > 1e-10 * ref["state"][:, :nsp].sum(axis=1, keepdims=True)
    big[:, nsp] = True
    err = np.where(big, rel, 0.0)[done].max(axis=1)
    assert np.mean(err <= 1e-6) >= 0.97


@pytest.mark.parametrize("nstrip", [1, 5, 64, 300])
def test_primordial_strip_be_bitwise(solver, prim_tables, nstrip):
    _, num, _ = primordial_ics(nstrip, seed=60 + nstrip)
    u = np.ascontiguousarray(num)
    rt = np.full_like(u, 1e-6)
    orc = oracle.StripSolver(prim_tables[1], (1.0, 1e9))
    solver.set_current_z(-1.0)
    for dt in (1e4, 1e8, 1e11):
        ug, ag = u.copy(), u * 1e-6
        uo, ao = u.copy(), u * 1e-6
        rg = solver.be_chem_solve(ug, dt, rt, ag, u.copy())
        ro = orc.be_chem_solve(uo, dt, rt, ao, u.copy())
        assert rg == ro
        assert np.array_equal(ag, ao) and np.array_equal(ug, uo)
        assert np.array_equal(solver.strip_temperatures(nstrip),
                              orc.strip_temperatures(nstrip))


def test_primordial_strip_failure_bitwise(solver, prim_tables):
    """A negative species fails the strip's RHS; the state comes back as the
    reference leaves it (multiplied by the scaling once more)."""
    _, num, _ = primordial_ics(8, seed=91)
    u = np.ascontiguousarray(num)
    sc = u.copy()
    u[5, 3] = -1e-30
    rt = np.full_like(u, 1e-6)
    orc = oracle.StripSolver(prim_tables[1], (1.0, 1e9))
    solver.set_current_z(-1.0)
    ug, ag = u.copy(), np.abs(u) * 1e-6
    uo, ao = u.copy(), np.abs(u) * 1e-6
    assert solver.be_chem_solve(ug, 1e8, rt, ag, sc.copy()) == 1
    assert orc.be_chem_solve(uo, 1e8, rt, ao, sc.copy()) == 1
    assert np.array_equal(ug, uo) and np.array_equal(ag, ao)


def test_primordial_run_harness_bitwise(solver, prim_tables):
    _, num, _ = primordial_ics(6, seed=77)
    ics = {s: num[:, j] * P_WEIGHTS[s] for j, s in enumerate(PRIMORDIAL_SPECIES)}
    args = (list(PRIMORDIAL_SPECIES), P_WEIGHTS, "ge", ics, 1e11)
    rv_g, t_g = harness.run_network(solver, *args, niter=600)
    rv_o, t_o = harness.run_network(oracle.StripSolver(prim_tables[1], (1.0, 1e9)), *args,
                                    niter=600)
    for k in rv_o:
        assert np.array_equal(rv_g[k], rv_o[k]), k
    for k in t_o:
        assert np.array_equal(t_g[k], t_o[k]), k
    assert t_o["successful"].any()


def test_carbon_run_harness_bitwise(carbon):
    mod, s = carbon
    _, n, _ = ion_ics(5, 43, 7, C_WEIGHT)
    ics = {sp: n[:, j] * mod.WEIGHTS[sp] for j, sp in enumerate(mod.SPECIES)}
    rv_g, t_g = mod.run_carbon(ics, 1e11, niter=500, solver=s)
    orc = oracle.StripSolver(s.tables, s.T_bounds, net=C_NET)
    rv_o, t_o = harness.run_network(orc, mod.SPECIES, mod.WEIGHTS, mod.ENERGY, ics, 1e11,
                                    niter=500)
    for k in rv_o:
        assert np.array_equal(rv_g[k], rv_o[k]), k
    for k in t_o:
        assert np.array_equal(t_g[k], t_o[k]), k
    assert t_o["successful"].any()
