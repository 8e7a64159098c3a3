"""Files of a solver: rate tables, initial conditions and solutions.

<name>_tables.bin, the rate-table file of this backend:

Layout (little endian): a 64-byte header -- magic "DENGOTB1", int32 number
of tables, int32 number of temperature bins nT, FP64 T0, T1, z0, z1, zero
padding -- then FP64 [ntables][nT], tables in the <name>_data struct order
(reactions, then cooling tables; dengo/chemical_network.py:310-318 writes
the same tables as HDF5 datasets /<rxn> and /<action>_<table> of
<name>_tables.h5).  The C side reads it in <name>_setup_data and
<name>_read_rate_tables / <name>_read_cooling_tables.

The reference's own HDF5 layouts are read and written through the HIP
library's HDF5 entry points (libhdf5_hl loaded at run time): <name>_tables.h5
(one FP64 dataset per table, /<rxn> and /<action>_<table>,
dengo/chemical_network.py:310-318), <name>_initial_conditions.h5 (one dataset
per species, :320-328) and <name>_solution.h5 (one dataset per species, /T,
root attributes time and timestep; rates_and_rate_tables.C.template:116-141).
Without HDF5 the same fields go to DENGOIC1 files: a 40-byte header (magic,
int32 nfields, int32 0, int64 ncells, FP64 time, FP64 timestep), nfields
names of 32 bytes, then FP64 [nfields][ncells].
"""
import ctypes

import numpy as np

MAGIC = b"DENGOTB1"
HEADER = 64


def write_blob(path, block, T_bounds, z_bounds=(0.0, 0.0)):
    block = np.ascontiguousarray(block, dtype="<f8")
    hdr = bytearray(HEADER)
    hdr[0:8] = MAGIC
    hdr[8:16] = np.array([block.shape[0], block.shape[1]], dtype="<i4").tobytes()
    hdr[16:48] = np.array([T_bounds[0], T_bounds[1], z_bounds[0], z_bounds[1]],
                          dtype="<f8").tobytes()
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(block.tobytes())


def read_blob(path):
    """(tables [ntables][nT], (T0, T1), (z0, z1)) of a table file."""
    with open(path, "rb") as f:
        hdr = f.read(HEADER)
        if len(hdr) != HEADER or hdr[:8] != MAGIC:
            raise ValueError("%s is not a DENGOTB1 table file" % path)
        nt, nT = np.frombuffer(hdr[8:16], dtype="<i4")
        b = np.frombuffer(hdr[16:48], dtype="<f8")
        data = np.frombuffer(f.read(), dtype="<f8")
    if data.size != int(nt) * int(nT):
        raise ValueError("%s: expected %d x %d values, found %d" % (path, nt, nT, data.size))
    return data.reshape(int(nt), int(nT)).copy(), (float(b[0]), float(b[1])), (float(b[2]),
                                                                              float(b[3]))


FIELDS_MAGIC = b"DENGOIC1"


def write_fields(path, fields, time=0.0, timestep=0.0):
    """Writes {name: 1-D array} (equal lengths, insertion order) as DENGOIC1."""
    names = list(fields)
    cols = [np.ascontiguousarray(fields[k], dtype="<f8").ravel() for k in names]
    n = cols[0].size if cols else 0
    if any(c.size != n for c in cols):
        raise ValueError("fields differ in length")
    with open(path, "wb") as f:
        f.write(FIELDS_MAGIC)
        f.write(np.array([len(names), 0], dtype="<i4").tobytes())
        f.write(np.array([n], dtype="<i8").tobytes())
        f.write(np.array([time, timestep], dtype="<f8").tobytes())
        for k in names:
            b = k.encode()
            if len(b) > 31:
                raise ValueError("field name %r longer than 31 bytes" % k)
            f.write(b + b"\0" * (32 - len(b)))
        for c in cols:
            f.write(c.tobytes())


def read_fields(path):
    """({name: array}, time, timestep) of a DENGOIC1 file."""
    with open(path, "rb") as f:
        hdr = f.read(40)
        if len(hdr) != 40 or hdr[:8] != FIELDS_MAGIC:
            raise ValueError("%s is not a DENGOIC1 field file" % path)
        nf = int(np.frombuffer(hdr[8:12], dtype="<i4")[0])
        n = int(np.frombuffer(hdr[16:24], dtype="<i8")[0])
        t, dt = np.frombuffer(hdr[24:40], dtype="<f8")
        names = [f.read(32).split(b"\0", 1)[0].decode() for _ in range(nf)]
        data = np.frombuffer(f.read(), dtype="<f8")
    if data.size != nf * n:
        raise ValueError("%s: expected %d x %d values, found %d" % (path, nf, n, data.size))
    data = data.reshape(nf, n)
    return {k: data[i].copy() for i, k in enumerate(names)}, float(t), float(dt)


def _h5():
    from . import runtime
    return runtime.load()


def hdf5_available():
    """True when the HIP library could load libhdf5_hl."""
    return bool(_h5().dengo_hdf5_available())


def _check(rc, what):
    if rc < 0:
        from . import runtime
        raise RuntimeError("%s: %s" % (what, runtime.last_error()))
    return rc


def write_h5(path, datasets, attrs=None):
    """Creates an HDF5 file of 1-D FP64 datasets {name: array} (names without
    a leading '/') and FP64 root attributes {name: value or 1-D array}."""
    lib = _h5()
    f = _check(lib.dengo_h5_open(path.encode(), 1), "dengo_h5_open")
    try:
        for k, v in datasets.items():
            a = np.ascontiguousarray(v, dtype=np.float64).ravel()
            _check(lib.dengo_h5_write(f, ("/" + k).encode(),
                                      a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.size),
                   "dengo_h5_write %s" % k)
        for k, v in (attrs or {}).items():
            if np.ndim(v) == 0:
                _check(lib.dengo_h5_set_attribute(f, k.encode(), float(v)), "attribute %s" % k)
            else:
                a = np.ascontiguousarray(v, dtype=np.float64).ravel()
                _check(lib.dengo_h5_set_attribute_array(
                    f, k.encode(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.size),
                    "attribute %s" % k)
    finally:
        lib.dengo_h5_close(f)


def read_h5(path, names, attrs=()):
    """({name: array}, {attr: value}) of the named datasets and root
    attributes of an HDF5 file (an attribute of more than one value comes back
    as an array)."""
    lib = _h5()
    f = _check(lib.dengo_h5_open(path.encode(), 0), "dengo_h5_open")
    out, at = {}, {}
    try:
        for k in names:
            n = _check(lib.dengo_h5_length(f, ("/" + k).encode()), "dataset %s" % k)
            a = np.empty(n, dtype=np.float64)
            _check(lib.dengo_h5_read(f, ("/" + k).encode(),
                                     a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n),
                   "dengo_h5_read %s" % k)
            out[k] = a
        for k in attrs:
            n = _check(lib.dengo_h5_attribute_length(f, k.encode()), "attribute %s" % k)
            v = np.empty(n, dtype=np.float64)
            _check(lib.dengo_h5_get_attribute_array(
                f, k.encode(), v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n),
                "attribute %s" % k)
            at[k] = float(v[0]) if n == 1 else v
    finally:
        lib.dengo_h5_close(f)
    return out, at


def write_tables_h5(path, table_names, block):
    """<name>_tables.h5 in the reference's layout: dataset /<table> per row of
    block, table_names as ChemicalNetwork.table_names() gives them."""
    write_h5(path, {k: block[i] for i, k in enumerate(table_names)})
