"""Strip semantics of the reference's Newton solver, on the CPU: the oracle's
strip BE_chem_solve (several cells sharing dt and the convergence decision,
as the Cython harness calls it) against the reference's own compiled
BE_chem_solve.C, and the run_<name> harness loop driven by both."""
import numpy as np
import pytest

import oracle
from dengo_amd import harness
from dengo_amd.ics import primordial_ics
from dengo_amd.networks import PRIMORDIAL_SPECIES, primordial_tables

WEIGHTS = dict(zip(PRIMORDIAL_SPECIES, [2.0, 2.0, 1.00794, 1.00794, 1.00794, 4.002602,
                                        4.002602, 4.002602, 1.0, 1.0]))


@pytest.fixture(scope="module")
def tab():
    return primordial_tables()[1]


def strip_state(n, seed):
    _, num, _ = primordial_ics(n, seed=seed)
    return np.ascontiguousarray(num)


def test_strip_electrons_run_across_cells():
    u = strip_state(5, 3)
    v = u.copy()
    oracle.StripSolver(np.zeros((45, 1024)), (1.0, 1e9)).ensure_electron_consistency(v)
    fe = np.array([0.0, 1.0, 0.0, 1.0, -1.0, 0.0, 1.0, 2.0, 0.0, 0.0])
    per_cell = (u * fe).sum(axis=1)
    assert np.allclose(v[:, 8], np.cumsum(per_cell), rtol=1e-14, atol=0)
    assert np.array_equal(np.delete(v, 8, axis=1), np.delete(u, 8, axis=1))


@pytest.mark.parametrize("nstrip", [1, 3, 8])
def test_strip_be_matches_reference(tab, nstrip):
    if oracle.load_ref() is None:
        pytest.skip("reference BE_chem_solve not built")
    u = strip_state(nstrip, 17)
    rt = np.full_like(u, 1e-6)
    at = u * 1e-6
    for dt in (1e6, 1e9, 1e12):
        out = []
        for use_ref in (False, True):
            s = oracle.StripSolver(tab, (1.0, 1e9), use_ref=use_ref)
            uu, aa = u.copy(), at.copy()
            rv = s.be_chem_solve(uu, dt, rt, aa, u.copy())
            out.append((rv, uu, aa, s.strip_temperatures(nstrip)))
        (r0, u0, a0, T0), (r1, u1, a1, T1) = out
        assert r0 == r1
        assert np.array_equal(u0, u1) and np.array_equal(a0, a1) and np.array_equal(T0, T1)


def test_strip_failure_state_matches_reference(tab):
    """A negative species in one cell fails the whole strip's RHS: every
    cell's state comes back multiplied by its scaling once more, as the
    reference leaves it."""
    if oracle.load_ref() is None:
        pytest.skip("reference BE_chem_solve not built")
    u = strip_state(4, 19)
    sc = u.copy()
    u[2, 3] = -1e-30
    rt = np.full_like(u, 1e-6)
    out = []
    for use_ref in (False, True):
        s = oracle.StripSolver(tab, (1.0, 1e9), use_ref=use_ref)
        uu, aa = u.copy(), np.abs(u) * 1e-6
        rv = s.be_chem_solve(uu, 1e8, rt, aa, sc.copy())
        out.append((rv, uu, aa))
    assert out[0][0] == out[1][0] == 1
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
    assert not np.array_equal(out[0][1], u)


def test_strip_couples_cells(tab):
    """A strip takes the sweeps its slowest cell needs: cells solved in one
    strip differ from the same cells solved one at a time."""
    u = strip_state(6, 5)
    rt = np.full_like(u, 1e-6)
    s = oracle.StripSolver(tab, (1.0, 1e9))
    a = u.copy()
    s.be_chem_solve(a, 1e10, rt, u * 1e-6, u.copy())
    b = u.copy()
    for i in range(6):
        bi = b[i:i + 1].copy()
        s.be_chem_solve(bi, 1e10, rt[i:i + 1], u[i:i + 1] * 1e-6, u[i:i + 1].copy())
        b[i] = bi[0]
    assert not np.array_equal(a, b)


def test_run_harness_reference_vs_oracle(tab):
    if oracle.load_ref() is None:
        pytest.skip("reference BE_chem_solve not built")
    num = strip_state(4, 23)
    ics = {s: num[:, j] * WEIGHTS[s] for j, s in enumerate(PRIMORDIAL_SPECIES)}
    res = []
    for use_ref in (False, True):
        s = oracle.StripSolver(tab, (1.0, 1e9), use_ref=use_ref)
        res.append(harness.run_network(s, list(PRIMORDIAL_SPECIES), WEIGHTS, "ge", ics, 1e10,
                                       niter=400, intermediate=1))
    (rv0, t0), (rv1, t1) = res
    for k in rv0:
        assert np.array_equal(rv0[k], rv1[k]), k
    for k in t0:
        assert np.array_equal(t0[k], t1[k]), k
    ok = t0["successful"]
    assert ok.any()
    # accepted steps advance t by dt; the clock never passes tf
    steps = np.nonzero(ok)[0]
    assert np.all(t0["t"][steps[1:]] >= t0["t"][steps[:-1]])
    assert t0["t"].max() <= 1e10
