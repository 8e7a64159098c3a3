"""Pins of the oracle's synthetic-network path (orc_big_*): portable rates
against the closed Arrhenius form, RHS invariants and a hand-computed
network, the Jacobian against finite differences, the BE integration against
the closed form of a linear decay, and the committed fixture against a fresh
oracle run."""
import os

import numpy as np
import pytest

from dengo_amd.synthetic import SyntheticNetwork, synthetic_ics
from dengo_amd.ics import tolerances
import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


class TinyNet(object):
    """A hand-written network in the SyntheticNetwork interface."""

    def __init__(self, n, rxn, A, beta=None, gamma=None):
        self.n = n
        self.rxn = np.asarray(rxn, dtype=np.int32).reshape(-1, 4)
        self.nrxn = self.rxn.shape[0]
        self.A = np.asarray(A, dtype=np.float64)
        self.beta = np.zeros(self.nrxn) if beta is None else np.asarray(beta, dtype=np.float64)
        self.gamma = np.zeros(self.nrxn) if gamma is None else np.asarray(gamma, dtype=np.float64)


@pytest.fixture(scope="module")
def net():
    return SyntheticNetwork(seed=7)


def test_portable_rates_match_closed_form(net):
    T = np.exp(np.random.default_rng(3).uniform(np.log(10.0), np.log(1e4), 64))
    k = oracle.big_rates(net, T)
    ref = net.rates(T)   # numpy: A (T/300)^beta exp(-gamma/T)
    rel = np.abs(k - ref) / ref
    assert rel.max() < 1e-13, rel.max()


def test_portable_exp_against_libm():
    import ctypes
    lib = oracle.load()
    lib.orc_pexp.restype = ctypes.c_double
    lib.orc_pexp.argtypes = [ctypes.c_double]
    xs = np.concatenate([np.random.default_rng(4).uniform(-740.0, 709.0, 20000),
                         [0.0, 1.0, -1.0, 0.5 * np.log(2.0), 700.0, -700.0]])
    y = np.array([lib.orc_pexp(float(x)) for x in xs])
    ref = np.exp(xs)
    ulp = np.abs(y - ref) / np.spacing(ref)
    assert ulp.max() <= 1.0, ulp.max()
    assert lib.orc_pexp(0.0) == 1.0 and lib.orc_pexp(-800.0) == 0.0
    assert np.isinf(lib.orc_pexp(710.0)) and np.isnan(lib.orc_pexp(float("nan")))


def test_rhs_hand_computed():
    # X0 + X1 -> X2 + X2 and X2 + X2 -> X0 + X3 at T = 300 K (k = A exactly)
    tn = TinyNet(4, [[0, 1, 2, 2], [2, 2, 0, 3]], [2.0, 0.5])
    u = np.array([[3.0, 5.0, 7.0, 11.0]])
    f, J, nneg = oracle.big_eval(tn, np.array([300.0]), u)
    r1, r2 = 2.0 * 3.0 * 5.0, 0.5 * 7.0 * 7.0
    assert nneg == 0
    assert np.array_equal(f[0], [-r1 + r2, -r1, 2 * r1 - 2 * r2, r2])
    # J[col][row] = d f_row / d u_col
    Jr = np.zeros((4, 4))
    Jr[0] = [-2.0 * 5.0, -2.0 * 5.0, 2 * 2.0 * 5.0, 0.0]            # d/du0
    Jr[1] = [-2.0 * 3.0, -2.0 * 3.0, 2 * 2.0 * 3.0, 0.0]            # d/du1
    Jr[2] = [2 * 0.5 * 7.0, 0.0, -2 * 2 * 0.5 * 7.0, 2 * 0.5 * 7.0]  # d/du2 (u2 squared)
    assert np.array_equal(J[0], Jr)


def test_rhs_conserves_particles(net):
    x, T = synthetic_ics(net, 16, seed=21)
    f, _, nneg = oracle.big_eval(net, T, x)
    assert nneg == 0
    scale = np.abs(f).sum(axis=1)
    assert np.all(np.abs(f.sum(axis=1)) <= 1e-13 * scale)


def test_negative_density_is_reported(net):
    x, T = synthetic_ics(net, 3, seed=22)
    x[1, 5] = -1e-30
    _, _, nneg = oracle.big_eval(net, T, x)
    assert nneg == 1


def test_jacobian_matches_finite_differences(net):
    x, T = synthetic_ics(net, 2, seed=23)
    f0, J, _ = oracle.big_eval(net, T, x)
    for c in range(2):
        for col in (0, 17, 64, 127):
            h = 1e-6 * x[c, col]
            xp = x[c:c + 1].copy()
            xm = x[c:c + 1].copy()
            xp[0, col] += h
            xm[0, col] -= h
            fp, _, _ = oracle.big_eval(net, T[c:c + 1], xp)
            fm, _, _ = oracle.big_eval(net, T[c:c + 1], xm)
            fd = (fp[0] - fm[0]) / (2 * h)
            # the RHS is bilinear: central differences are exact up to rounding
            scale = np.abs(f0[c]).max() / x[c, col]
            assert np.abs(fd - J[c, col]).max() <= 1e-6 * scale + 1e-300


def test_be_linear_decay_closed_form():
    # X0 + X1 -> X1 + X2: X1 is a catalyst, so X0 decays with rate k u1 and the
    # BE step is u0 / (1 + dt k u1) exactly; the substep sequence is the
    # reference's dt0 = dtf/1e5, dt <- min(1.1 dt, dtf - t).
    k, u1 = 3e-9, 2.0e3
    tn = TinyNet(3, [[0, 1, 1, 2]], [k])
    x = np.array([[1.0e4, u1, 1.0]])
    rtol = np.full_like(x, 1e-9)
    atol = 1e-9 * x
    dtf = 1e7
    xs = x.copy()
    ttot, _, out = oracle.big_evolve(tn, xs, np.array([300.0]), rtol, atol.copy(), dtf)
    dt, t, u0 = dtf / 1e5, 0.0, x[0, 0]
    n = 0
    while t < dtf:
        u0 = u0 / (1.0 + dt * k * u1)
        t += dt
        dt = min(dt * 1.1, dtf - t)
        n += 1
    assert out["status"][0] == 0 and out["nsub"][0] == n and out["nrej"][0] == 0
    assert abs(xs[0, 0] - u0) <= 1e-9 * u0
    assert xs[0, 1] == u1
    assert abs(xs[0].sum() - x[0].sum()) <= 1e-12 * x[0].sum()
    assert ttot == t


def test_zero_rates_step_count():
    # f = 0: every substep is accepted in one sweep; Sum_{k<n} dt0 1.1^k >= dtf
    tn = TinyNet(2, [[0, 1, 1, 1]], [0.0])
    x = np.array([[2.0, 3.0]])
    xs = x.copy()
    _, _, out = oracle.big_evolve(tn, xs, np.array([300.0]), np.full_like(x, 1e-9), 1e-9 * x, 1e12)
    assert out["nsub"][0] == 97 and np.array_equal(xs, x)


def test_fixture_matches_fresh_oracle_run():
    d = np.load(os.path.join(HERE, "golden", "synthetic_cells.npz"))
    net = SyntheticNetwork(seed=int(d["net_seed"]))
    x = d["input"].copy()
    at = d["atol"].copy()
    sel = [0, 3, 7, 16]   # quick cells
    xs = np.ascontiguousarray(x[sel])
    ats = np.ascontiguousarray(at[sel])
    _, _, out = oracle.big_evolve(net, xs, d["T"][sel], d["rtol"][sel], ats, float(d["dtf"]))
    assert np.array_equal(xs, d["state"][sel])
    assert np.array_equal(ats, d["atol_out"][sel])
    assert np.array_equal(out["nsweep"], d["nsweep"][sel])
    x2, T2 = synthetic_ics(net, 24, seed=int(d["seed"]))
    assert np.array_equal(x2, d["input"]) and np.array_equal(T2, d["T"])
    rtol, atol = tolerances(x2)
    assert np.array_equal(rtol, d["rtol"]) and np.array_equal(atol, d["atol"])


def test_substep_cap_does_not_leak_between_calls():
    # the oracle's substep cap is a library-wide setting: a call with a small
    # cap must not cap the next call that asks for the reference's default
    from dengo_amd.ics import primordial_ics, tolerances
    from dengo_amd.networks import primordial_tables
    _, tab = primordial_tables()
    inp, _, _ = primordial_ics(2, seed=5)
    rt, at = tolerances(inp)
    r = oracle.prim_evolve(inp, rt, at, tab, 1e12, max_substeps=3)
    assert (r["nsub"] <= 3).all()
    tn = TinyNet(2, [[0, 1, 1, 1]], [0.0])
    x = np.array([[2.0, 3.0]])
    _, _, out = oracle.big_evolve(tn, x, np.array([300.0]), np.full_like(x, 1e-9), 1e-9 * x, 1e12)
    assert out["status"][0] == 0 and out["nsub"][0] == 97
