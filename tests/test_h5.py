"""HDF5 files in the reference's layouts, on the CPU: the library's HDF5 entry
points round-trip datasets and attributes, <name>_setup_data reads the
reference's <name>_tables.h5 (dengo/chemical_network.py:310-318) into the same
struct as the binary table file, and <name>_main (rates_and_rate_tables.C.
template:56-147) reports a missing initial-conditions file as the reference
does (return 1).  Skipped where libhdf5_hl cannot be loaded."""
import ctypes
import os

import numpy as np
import pytest

from dengo_amd import runtime, tables
from dengo_amd.networks import oxygen_network, primordial_network

pytestmark = pytest.mark.skipif(not tables.hdf5_available(), reason="libhdf5_hl not found")


def test_datasets_and_attributes_round_trip(tmp_path):
    p = str(tmp_path / "x.h5")
    rng = np.random.default_rng(5)
    data = {"a": rng.standard_normal(17), "b_c": np.array([1e-300, -0.0, np.inf, 3.5])}
    tables.write_h5(p, data, attrs={"time": 1.25e12, "timestep": 7.0})
    assert runtime.load().dengo_h5_is_hdf5(p.encode()) == 1
    got, at = tables.read_h5(p, ["a", "b_c"], attrs=["time", "timestep"])
    for k in data:
        assert np.array_equal(got[k], data[k]) and got[k].dtype == np.float64
    assert at == {"time": 1.25e12, "timestep": 7.0}


def test_array_attributes_round_trip(tmp_path):
    """Per-cell root attributes (the intermediate files' cell_time and
    cell_timestep) beside scalar ones; reading checks the stored length."""
    p = str(tmp_path / "x.h5")
    ct = np.array([1.0, 2.5e10, np.nan, 1e12])
    tables.write_h5(p, {"a": np.arange(4.0)}, attrs={"time": 1.0, "cell_time": ct})
    _, at = tables.read_h5(p, [], attrs=["time", "cell_time"])
    assert at["time"] == 1.0
    assert np.array_equal(at["cell_time"], ct, equal_nan=True)
    lib = runtime.load()
    f = lib.dengo_h5_open(p.encode(), 0)
    try:
        assert lib.dengo_h5_attribute_length(f, b"cell_time") == 4
        assert lib.dengo_h5_attribute_length(f, b"time") == 1
        v = ctypes.c_double()
        # a scalar read of a 4-value attribute is refused, not overrun
        assert lib.dengo_h5_get_attribute(f, b"cell_time", ctypes.byref(v)) < 0
        assert "holds 4 values, not 1" in runtime.last_error()
        assert lib.dengo_h5_attribute_length(f, b"nope") < 0
    finally:
        lib.dengo_h5_close(f)


def test_missing_dataset_and_wrong_length_are_errors(tmp_path):
    p = str(tmp_path / "x.h5")
    tables.write_h5(p, {"a": np.arange(4.0)})
    lib = runtime.load()
    f = lib.dengo_h5_open(p.encode(), 0)
    assert f >= 0
    try:
        assert lib.dengo_h5_length(f, b"/nope") < 0
        assert "no dataset" in runtime.last_error()
        buf = np.zeros(5)
        P = ctypes.POINTER(ctypes.c_double)
        assert lib.dengo_h5_read(f, b"/a", buf.ctypes.data_as(P), 5) < 0
        assert "holds 4 values, expected 5" in runtime.last_error()
    finally:
        lib.dengo_h5_close(f)
    assert lib.dengo_h5_open(str(tmp_path / "absent.h5").encode(), 0) < 0
    assert lib.dengo_h5_is_hdf5(str(tmp_path / "absent.h5").encode()) == 0


def _struct_bytes(lib, name, ptr):
    """The struct's bytes up to its last field, device_state (a pointer to
    this process's per-struct state, different for each struct)."""
    n = getattr(lib, "%s_data_bytes" % name)()
    return ctypes.string_at(ptr, n - ctypes.sizeof(ctypes.c_void_p))


@pytest.mark.parametrize("name", ["primordial", "oxygen"])
def test_setup_data_reads_reference_table_file(tmp_path, name):
    with np.errstate(all="ignore"):
        net = primordial_network() if name == "primordial" else oxygen_network()
    net.write_tables(name, str(tmp_path))
    h5 = net.write_tables_h5(name, str(tmp_path))
    # the file holds the reference's dataset names, one per table
    got, _ = tables.read_h5(h5, net.table_names())
    block = net.table_block()
    for i, k in enumerate(net.table_names()):
        assert np.array_equal(got[k], block[i]), k
    lib = runtime.load()
    setup = getattr(lib, "%s_setup_data_from" % name)
    free = getattr(lib, "%s_free_data" % name)
    a = setup(str(tmp_path / ("%s_tables.bin" % name)).encode(), None, None)
    b = setup(h5.encode(), None, None)
    assert a and b
    try:
        # same tables, grid and (zeroed) scratch: the structs are byte-identical
        assert _struct_bytes(lib, name, a) == _struct_bytes(lib, name, b)
    finally:
        free(a)
        free(b)


def test_setup_data_prefers_the_reference_file_in_cwd(tmp_path, monkeypatch):
    with np.errstate(all="ignore"):
        net = primordial_network()
    net.write_tables_h5("primordial", str(tmp_path))
    monkeypatch.chdir(tmp_path)
    lib = runtime.load()
    # only the .h5 exists: <n>_setup_data reads it (rates_and_rate_tables.C.template:280)
    d = lib.primordial_setup_data(None, None)
    assert d
    lib.primordial_free_data(d)
    os.remove("primordial_tables.h5")
    assert not lib.primordial_setup_data(None, None)


def test_main_without_initial_conditions_returns_1(tmp_path, monkeypatch):
    with np.errstate(all="ignore"):
        net = primordial_network()
    net.write_tables("primordial", str(tmp_path))
    monkeypatch.chdir(tmp_path)
    assert runtime.load().primordial_main(0, None) == 1


def test_field_files_round_trip(tmp_path):
    p = str(tmp_path / "f.bin")
    fields = {"H2I": np.arange(3.0), "ge": np.array([1.0, 2.0, 3.0])}
    tables.write_fields(p, fields, time=2.0, timestep=0.5)
    got, t, dt = tables.read_fields(p)
    assert list(got) == ["H2I", "ge"] and t == 2.0 and dt == 0.5
    for k in fields:
        assert np.array_equal(got[k], fields[k])
