"""The rate-table file of a solver: <name>_tables.bin.

Layout (little endian): a 64-byte header -- magic "DENGOTB1", int32 number
of tables, int32 number of temperature bins nT, FP64 T0, T1, z0, z1, zero
padding -- then FP64 [ntables][nT], tables in the <name>_data struct order
(reactions, then cooling tables; dengo/chemical_network.py:310-318 writes
the same tables as HDF5 datasets /<rxn> and /<action>_<table> of
<name>_tables.h5).  The C side reads it in <name>_setup_data and
<name>_read_rate_tables / <name>_read_cooling_tables.
"""
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
