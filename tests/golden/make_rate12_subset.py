"""Writes tests/golden/rate12_hhe.txt: the lines of the reference's UMIST
table (dengo/RATE12.txt) whose reactants and products all lie in a small
hydrogen-helium species set, unchanged (data for the RATE12 loader tests).

    python tests/golden/make_rate12_subset.py [/root/reference]
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SPECIES = {"H", "H2", "H+", "H2+", "H3+", "H-", "e-", "He", "He+", "HeH+"}


def main(ref):
    keep = []
    with open(os.path.join(ref, "dengo", "RATE12.txt")) as f:
        for line in f:
            r = line.rstrip("\n").split(":")
            if all(x in SPECIES for x in r[2:6]) and all(x == "" or x in SPECIES for x in r[6:8]):
                keep.append(line)
    with open(os.path.join(HERE, "rate12_hhe.txt"), "w") as f:
        f.writelines(keep)
    print(len(keep), "lines")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
