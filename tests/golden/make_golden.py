"""Regenerate the primordial golden fixtures from the CPU oracle.

Writes tests/golden/primordial_cells.npz: 96 seeded cells (config-1
distributions, seed 11) with inputs, tolerances and the oracle's outputs
(final state, atol, elapsed time, per-cell counters) for ln T taken with libm
("libm_*") and with the portable log the device code uses ("plog_*"), and
tests/golden/oxygen_cells.npz: the same for 128 O I-IX cells (seed 12) of the
synthetic oxygen network (table seed 3), and
tests/golden/synthetic_cells.npz: 24 cells (seed 13) of the synthetic
128-species network (reaction seed 7), whose oracle uses the portable log
and exp for its rates.
dtf = 1e12 s, dt = -1, z = -1, as the emitted main runs the solver.
Usage: make_golden.py [primordial|oxygen|synthetic ...] (default: all).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
from dengo_amd.ics import primordial_ics, tolerances  # noqa: E402
from dengo_amd.ics import oxygen_ics  # noqa: E402
from dengo_amd.networks import oxygen_tables, primordial_tables  # noqa: E402
from dengo_amd.synthetic import SyntheticNetwork, synthetic_ics  # noqa: E402


def primordial():
    _, tab = primordial_tables()
    inp, _, _ = primordial_ics(96, seed=11)
    rtol, atol = tolerances(inp)
    out = dict(input=inp, rtol=rtol, atol=atol, dtf=1e12, seed=11)
    for tag, flags in (("libm", 0), ("plog", oracle.PORTABLE_LOG)):
        r = oracle.prim_evolve(inp, rtol, atol, tab, 1e12, flags=flags)
        for k in ("state", "atol", "status", "nsub", "nrej", "nsweep", "dt_cell", "T"):
            out["%s_%s" % (tag, k)] = r[k]
        out["%s_ttot" % tag] = r["ttot"]
        out["%s_dt" % tag] = r["dt"]
    np.savez_compressed(os.path.join(HERE, "primordial_cells.npz"), **out)
    print("wrote primordial_cells.npz; capped cells:", int((out["libm_status"] == 2).sum()))


def oxygen():
    _, otab = oxygen_tables()
    net = oracle.ion_chain_net(8, 15.9994)
    inp, _, _ = oxygen_ics(128, seed=12)
    rtol, atol = tolerances(inp)
    out = dict(input=inp, rtol=rtol, atol=atol, dtf=1e12, seed=12)
    for tag, flags in (("libm", 0), ("plog", oracle.PORTABLE_LOG)):
        r = oracle.ma_evolve(net, inp, rtol, atol, otab, 1e12, (1e4, 1e9), flags=flags)
        for k in ("state", "atol", "status", "nsub", "nrej", "nsweep", "dt_cell", "T"):
            out["%s_%s" % (tag, k)] = r[k]
        out["%s_ttot" % tag] = r["ttot"]
        out["%s_dt" % tag] = r["dt"]
    np.savez_compressed(os.path.join(HERE, "oxygen_cells.npz"), **out)
    print("wrote oxygen_cells.npz")


def synthetic():
    net = SyntheticNetwork(seed=7)
    x, T = synthetic_ics(net, 24, seed=13)
    rtol, atol = tolerances(x)
    out = dict(input=x.copy(), T=T, rtol=rtol, atol=atol.copy(), dtf=1e12, seed=13, net_seed=7)
    xs = x.copy()
    ttot, dt, r = oracle.big_evolve(net, xs, T, rtol, atol, 1e12)
    out.update(state=xs, atol_out=atol, ttot=ttot, dt=dt, status=r["status"], nsub=r["nsub"],
               nrej=r["nrej"], nsweep=r["nsweep"], dt_cell=r["dt"])
    np.savez_compressed(os.path.join(HERE, "synthetic_cells.npz"), **out)
    print("wrote synthetic_cells.npz; sweeps per cell:", r["nsweep"].tolist())


if __name__ == "__main__":
    which = sys.argv[1:] or ["primordial", "oxygen", "synthetic"]
    for w in which:
        {"primordial": primordial, "oxygen": oxygen, "synthetic": synthetic}[w]()
