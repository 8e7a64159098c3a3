"""Pins on the oracle's intermediate-solution records (no GPU).

The reference's dengo_evolve writes, after every accepted substep, the state
before that substep (its `prev`, number densities, not yet floored), the
temperatures of the last Newton sweep, the elapsed time and the next dt
(rates_and_rate_tables.C.template:199-241).  The checks below tie the records
to runs that do not use the recording code:

* record 0 is the entry state: mass densities divided by weight * 1.67e-24 and
  the electron pass (:158-170, :604-630), computed in numpy;
* record k + 1 is the final state of the same cell run with the accepted-
  substep cap at k: that run stops at its (k + 1)-th acceptance before the
  state is floored or converted (:203-204), so it ends on the state the
  uncapped run's record k + 1 shows as `prev`;
* record k - 1's elapsed time is that capped run's elapsed time, and record
  k's temperature that run's last-sweep temperature;
* elapsed times grow, and the next dt never exceeds dtf minus the elapsed
  time (zero once the cell reaches dtf).
"""
import numpy as np
import pytest

import oracle
from dengo_amd.ics import primordial_ics, tolerances
from dengo_amd.networks import primordial_tables

W = np.array([2.0, 2.0, 1.00794, 1.00794, 1.00794, 4.002602, 4.002602, 4.002602, 1.0, 1.0])
FREE_E = [0.0, 1.0, 0.0, 1.0, -1.0, 0.0, 1.0, 2.0]
N = 10
CAP = 400
DTF = 1e12


@pytest.fixture(scope="module")
def run():
    tables = primordial_tables()[1]
    x, _, _ = primordial_ics(12, seed=123, T_range=(50.0, 1e6))
    rt, at = tolerances(x)
    recs = []
    for c in range(x.shape[0]):
        with oracle.trajectory(1, N, CAP) as buf:
            r = oracle.prim_evolve(x[c:c + 1], rt[c:c + 1], at[c:c + 1], tables, DTF,
                                   max_substeps=CAP)
        recs.append(buf[:, 0, :].copy())
        assert r["nsub"][0] >= 1
    return tables, x, rt, at, recs


def _number(x):
    """The driver's entry conversion and electron pass, in numpy."""
    u = x.copy()
    for k in range(N - 1):
        u[:, k] = u[:, k] / (W[k] * 1.67e-24)
    for c in range(u.shape[0]):
        tot = 0.0
        for k in range(8):
            tot += u[c, k] * FREE_E[k]
        u[c, 8] = tot
    return u


def test_first_record_is_the_entry_state(run):
    _, x, _, _, recs = run
    u = _number(x)
    for c, rec in enumerate(recs):
        assert np.array_equal(rec[0, :N], u[c]), c


def test_records_match_capped_runs(run):
    tables, x, rt, at, recs = run
    u = _number(x)
    atn = at.copy()
    for k in range(N - 1):
        atn[:, k] = atn[:, k] / (W[k] * 1.67e-24)
    checked = 0
    for c, rec in enumerate(recs):
        n_rec = int(np.sum(~np.isnan(rec[:, N + 1])))
        for k in sorted({1, 2, 5, n_rec // 2, n_rec - 2}):
            if k < 1 or k + 1 >= n_rec:
                continue
            r = oracle.prim_evolve(u[c:c + 1], rt[c:c + 1], atn[c:c + 1], tables, DTF,
                                   input_is_number=True, max_substeps=k)
            assert r["status"][0] == 2
            assert np.array_equal(r["state"][0], rec[k + 1, :N]), (c, k)
            assert r["ttot"] == rec[k - 1, N + 1], (c, k)
            assert r["T"][0] == rec[k, N], (c, k)
            checked += 1
    assert checked >= 20


def test_times_and_next_steps(run):
    _, _, _, _, recs = run
    for rec in recs:
        ok = ~np.isnan(rec[:, N + 1])
        t, dt = rec[ok, N + 1], rec[ok, N + 2]
        assert np.all(np.diff(t) > 0)
        # the next dt is min(1.1 dt, dtf - t): zero once the cell reaches dtf
        assert np.all(dt <= DTF - t) and np.all((dt > 0) | (t == DTF))
        assert np.all(rec[ok, N] > 0)
