"""<name>_main on the GPU (rates_and_rate_tables.C.template:56-147): tables
and initial conditions from the working directory, in the reference's HDF5
layout or this backend's binary files, integrated to the network's stop time
with atol = 1e-9 x IC and rtol = 1e-9; the solution file must hold exactly
what dengo_evolve_<name> returns for the same arrays."""
import os

import numpy as np
import pytest

from dengo_amd import runtime, tables
from dengo_amd.ics import primordial_ics
from dengo_amd.networks import primordial_network

pytestmark = pytest.mark.gpu

FIELDS = ("H2I", "H2II", "HI", "HII", "HM", "HeI", "HeII", "HeIII", "de", "ge")


@pytest.mark.parametrize("fmt", ["h5", "bin"])
def test_main_matches_evolve(tmp_path, monkeypatch, fmt, solver):
    if fmt == "h5" and not tables.hdf5_available():
        pytest.skip("libhdf5_hl not found")
    with np.errstate(all="ignore"):
        net = primordial_network()
    inp, _, _ = primordial_ics(96, seed=31)
    ics = {k: inp[:, i].copy() for i, k in enumerate(FIELDS)}
    if fmt == "h5":
        net.write_tables_h5("primordial", str(tmp_path))
        tables.write_h5(str(tmp_path / "primordial_initial_conditions.h5"), ics)
    else:
        net.write_tables("primordial", str(tmp_path))
        tables.write_fields(str(tmp_path / "primordial_initial_conditions.bin"), ics)
    monkeypatch.chdir(tmp_path)
    print("primordial_main, %s files, 96 cells" % fmt, flush=True)
    assert runtime.load().primordial_main(0, None) == 0
    ref = solver.evolve(inp, np.full_like(inp, 1e-9), inp * 1e-9, net.stop_time)
    # the solution is written as HDF5 whenever libhdf5 loads (the reference's
    # file), whatever the input files were
    if tables.hdf5_available():
        got, at = tables.read_h5("primordial_solution.h5", list(FIELDS) + ["T"],
                                 attrs=["time", "timestep"])
        t, dt = at["time"], at["timestep"]
        assert not os.path.exists("primordial_solution.bin")
    else:
        got, t, dt = tables.read_fields("primordial_solution.bin")
        assert not os.path.exists("primordial_solution.h5")
    for i, k in enumerate(FIELDS):
        assert np.array_equal(got[k], ref["state"][:, i]), k
    assert np.array_equal(got["T"], ref["T"])
    assert t == ref["ttot"] and dt == ref["dt"]
