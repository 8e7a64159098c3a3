"""Writes the data fixtures of the data-file network builders: the oxygen
datasets of the reference's input files (Gnat & Ferland ion cooling on their
T grid, HM12 photoionisation and photoheating on their z grid), one .npz per
file under tests/golden/ion_data/input/<kind>/ with the files' dataset names,
so that dengo_amd.ion_by_ion can build the network where the reference tree
or HDF5 is absent.  Reads the HDF5 files through dengo_amd.tables.read_h5.

    python tests/golden/make_ion_data.py [/root/reference]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from dengo_amd.tables import read_h5  # noqa: E402

FILES = [
    ("cooling", "o_ion_by_ion_cooling", ["T"] + ["o_%d" % k for k in range(1, 10)]),
    ("photoionization", "o_ion_by_ion_photoionization_HM12", ["z"] + ["o_%d" % k for k in range(1, 9)]),
    ("photoheating", "o_ion_by_ion_photoheating_HM12", ["z"] + ["o_%d" % k for k in range(1, 9)]),
]


def main(ref):
    for kind, name, datasets in FILES:
        d, _ = read_h5(os.path.join(ref, "input", kind, name + ".h5"), datasets)
        out = os.path.join(HERE, "ion_data", "input", kind)
        os.makedirs(out, exist_ok=True)
        np.savez(os.path.join(out, name + ".npz"), **d)
        print(kind, name, {k: v.shape for k, v in d.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
