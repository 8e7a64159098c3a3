"""dengo_evolve_<n> with intermediate-solution files on the GPU
(rates_and_rate_tables.C.template:199-241, ChemicalNetwork(write_intermediate=
True)).

File k holds, per cell, what the reference writes after the k-th accepted
substep: the state before it (number densities), /T of the last Newton sweep,
and the elapsed time and next dt.  Every record is checked bit for bit
against the oracle's records of the same cells (tests/test_oracle_intermediate
pins those against capped runs); a cell with fewer than k substeps repeats its
last record; the scalar time/timestep attributes are those of the cell with the
least elapsed time.  While the files are on, the evolve runs the
reference-faithful sweep whatever precision is asked for."""
import glob
import os

import numpy as np
import pytest

import oracle
from dengo_amd import runtime, tables
from dengo_amd.ics import oxygen_ics, primordial_ics, tolerances
from dengo_amd.networks import oxygen_tables, primordial_tables

pytestmark = pytest.mark.gpu

PRIM_FIELDS = ["H2I", "H2II", "HI", "HII", "HM", "HeI", "HeII", "HeIII", "de", "ge"]
OXY_FIELDS = ["O_1", "O_2", "O_3", "O_4", "O_5", "O_6", "O_7", "O_8", "O_9", "de", "ge"]
CAP = 300
DTF = 1e12


def _expected(buf, nsub, k):
    """File k's columns from the oracle's records [cap, cells, N + 3]."""
    m = np.minimum(nsub, buf.shape[0])
    rows = np.full((buf.shape[1], buf.shape[2]), np.nan)
    have = m > 0
    idx = np.minimum(k, m)[have] - 1
    rows[have] = buf[idx, np.nonzero(have)[0]]
    return rows


def _read(name, k, fields, n):
    if tables.hdf5_available():
        path = "%s_intermediate_%06d.h5" % (name, k)
        d, a = tables.read_h5(path, fields + ["T"],
                              attrs=["time", "timestep", "cell_time", "cell_timestep"])
        ct, cdt = np.atleast_1d(a["cell_time"]), np.atleast_1d(a["cell_timestep"])
        assert ct.shape == (n,) and cdt.shape == (n,)
        return d, a["time"], a["timestep"], ct, cdt
    d, t, dt = tables.read_fields("%s_intermediate_%06d.bin" % (name, k))
    return d, t, dt, None, None


def _check_files(name, fields, buf, nsub, n):
    K = int(min(CAP, nsub.max()))
    ext = "h5" if tables.hdf5_available() else "bin"
    files = sorted(glob.glob("%s_intermediate_*.%s" % (name, ext)))
    assert len(files) == K
    assert os.path.basename(files[0]) == "%s_intermediate_000001.%s" % (name, ext)
    N = len(fields)
    for k in range(1, K + 1):
        want = _expected(buf, nsub, k)
        d, t, dt, ct, cdt = _read(name, k, fields, n)
        for i, f in enumerate(fields):
            assert np.array_equal(d[f], want[:, i], equal_nan=True), (k, f)
        assert np.array_equal(d["T"], want[:, N], equal_nan=True), k
        if ct is not None:
            assert np.array_equal(ct, want[:, N + 1], equal_nan=True), k
            assert np.array_equal(cdt, want[:, N + 2], equal_nan=True), k
        j = int(np.nanargmin(want[:, N + 1]))
        assert t == want[j, N + 1] and dt == want[j, N + 2], k


def test_intermediate_files_primordial(tmp_path, monkeypatch):
    print("intermediate files: primordial, 48 cells, cap %d" % CAP, flush=True)
    monkeypatch.chdir(tmp_path)
    tab = primordial_tables()[1]
    x, _, _ = primordial_ics(48, seed=321, T_range=(50.0, 1e6))
    rt, at = tolerances(x)
    s = runtime.PrimordialSolver(tab)
    try:
        s.set_write_intermediate(True)
        got = s.evolve(x, rt, at, DTF, max_substeps=CAP, precision=runtime.FP64_FAST)
    finally:
        s.close()
    with oracle.trajectory(x.shape[0], 10, CAP) as buf:
        ref = oracle.prim_evolve(x, rt, at, tab, DTF, flags=oracle.PORTABLE_LOG,
                                 max_substeps=CAP)
    for k in ("state", "status", "nsub", "nrej"):
        assert np.array_equal(got[k], ref[k]), k
    assert (ref["status"] == 2).any() and (ref["nsub"] < CAP).any()
    _check_files("primordial", PRIM_FIELDS, buf, ref["nsub"], x.shape[0])


def test_intermediate_files_oxygen_one_cell(tmp_path, monkeypatch):
    """One cell: the scalar attributes are the cell's own, as in the reference."""
    print("intermediate files: oxygen, one cell", flush=True)
    monkeypatch.chdir(tmp_path)
    tab = oxygen_tables()[1]
    x, _, _ = oxygen_ics(1, seed=322)
    rt, at = tolerances(x)
    s = runtime.OxygenSolver(tab)
    try:
        s.set_write_intermediate(True)
        got = s.evolve(x, rt, at, DTF, max_substeps=CAP)
    finally:
        s.close()
    net = oracle.ion_chain_net(8, 15.9994)
    with oracle.trajectory(1, 11, CAP) as buf:
        ref = oracle.ma_evolve(net, x, rt, at, tab, DTF, s.T_bounds, flags=oracle.PORTABLE_LOG,
                               max_substeps=CAP)
    assert np.array_equal(got["state"], ref["state"])
    _check_files("oxygen", OXY_FIELDS, buf, ref["nsub"], 1)


def test_off_writes_nothing(tmp_path, monkeypatch):
    print("intermediate files: off", flush=True)
    monkeypatch.chdir(tmp_path)
    x, _, _ = primordial_ics(8, seed=323)
    rt, at = tolerances(x)
    s = runtime.PrimordialSolver(primordial_tables()[1])
    try:
        s.set_write_intermediate(True)
        s.set_write_intermediate(False)
        s.evolve(x, rt, at, DTF, max_substeps=50)
    finally:
        s.close()
    assert not glob.glob("primordial_intermediate_*")
