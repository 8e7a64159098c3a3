"""The solver template set on the CPU: the committed network sources are the
generator's output, write_solver emits the full set for a new network, and
the built library exports the declared entry points."""
import os
import re

import numpy as np

from dengo_amd import codegen, runtime
from dengo_amd.networks import ion_network, oxygen_network, primordial_network
from dengo_amd.tables import hdf5_available, read_blob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dengo_amd", "csrc")
CARBON = os.path.join(ROOT, "tests", "generated", "carbon")


def _read(*p):
    with open(os.path.join(*p)) as f:
        return f.read()


def test_builtin_sources_are_generated():
    with np.errstate(all="ignore"):
        prim = primordial_network()
    oxy = oxygen_network()
    assert codegen.render_instantiation(prim, "primordial", kind="prim") == \
        _read(CSRC, "dengo_primordial.hip")
    assert codegen.render_instantiation(oxy, "oxygen") == _read(CSRC, "dengo_oxygen.hip")
    assert codegen.render_device_network(oxy, "oxygen") == _read(CSRC, "net_oxygen.h")
    assert codegen.render_data_header(oxy, "oxygen") == _read(ROOT, "include", "oxygen_solver.h")


def test_write_solver_emits_template_set(tmp_path):
    net = ion_network("C", seed=9)
    written = net.write_solver("carbon", output_dir=str(tmp_path))
    names = sorted(os.path.basename(p) for p in written)
    expect = ["carbon_solver.h", "net_carbon.h", "carbon_solver.hip", "carbon_run.py",
              "carbon_tables.bin"]
    if hdf5_available():
        expect.append("carbon_tables.h5")
    assert names == sorted(expect)
    for f in ("carbon_solver.h", "net_carbon.h", "carbon_solver.hip", "carbon_run.py",
              "carbon_tables.json"):
        assert _read(str(tmp_path), f) == _read(CARBON, f), f
    tab, Tb, _ = read_blob(str(tmp_path / "carbon_tables.bin"))
    assert tab.shape == (len(net.table_names()), 1024) and Tb == (1e4, 1e9)
    assert np.array_equal(tab, net.table_block())
    hdr = _read(str(tmp_path), "carbon_solver.h")
    assert "#define CARBON_NSPECIES 9" in hdr and "#define CARBON_NTABLES 19" in hdr


def test_generated_library_exports(tmp_path):
    path = os.path.join(CARBON, "libcarbon_solver.so")
    lib = runtime.load_solver_library(path, "carbon")
    hdr = re.sub(r"/\*.*?\*/", "", _read(CARBON, "carbon_solver.h"), flags=re.S)
    for name in re.findall(r"\b((?:dengo_)?\w*carbon\w*)\s*\(", hdr):
        assert hasattr(lib, name), name
