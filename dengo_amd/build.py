"""Build the HIP library in-tree: dengo_amd/lib/libdengo_hip.so (gfx950).

Flags: -ffp-contract=off keeps every FP64 expression rounding as written (the
reference-faithful kernels depend on it); packed FP32 instructions are not
generated.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libdengo_hip.so")
SOURCES = ["dengo_common.hip", "dengo_primordial.hip", "dengo_oxygen.hip", "dengo_synthetic.hip",
           "dengo_server.hip", "dengo_graph.hip", "dengo_h5.hip", "dengo_noalloc.hip",
           "dengo_probe.hip"]
HEADERS = ["dengo_device.h", "be_driver.h", "be_fast.h", "lean_sweep.h", "gen_net.h", "ma_net.h", "dengo_solver.h", "net_oxygen.h",
           "big_net.h", "big_program.h", "arena.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False, extra=()):
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra) + [
        "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-o", LIB + ".tmp"] + srcs + ["-ldl"]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed (%d)" % r.returncode)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_solver(output_dir, solver_name, verbose=False):
    """Compiles the solver ChemicalNetwork.write_solver wrote into output_dir
    (<name>_solver.hip and its headers) with this package's kernels into
    output_dir/lib<name>_solver.so."""
    src = os.path.join(output_dir, "%s_solver.hip" % solver_name)
    if not os.path.exists(src):
        raise RuntimeError("%s not found: run write_solver first" % src)
    lib = os.path.join(output_dir, "lib%s_solver.so" % solver_name)
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", output_dir, "-I", CSRC, "-I", os.path.join(ROOT, "include"),
                                     "-o", lib + ".tmp", src, os.path.join(CSRC, "dengo_common.hip"),
                                     os.path.join(CSRC, "dengo_h5.hip"), "-ldl"]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed (%d)" % r.returncode)
    os.replace(lib + ".tmp", lib)
    return lib


GENERATED = os.path.join(ROOT, "tests", "generated")
# the oxygen datasets of the reference's input files (tests/golden/make_ion_data.py)
ION_DATA = os.path.join(ROOT, "tests", "golden", "ion_data")
# the hydrogen-helium lines of the reference's UMIST table (tests/golden/make_rate12_subset.py)
RATE12_HHE = os.path.join(ROOT, "tests", "golden", "rate12_hhe.txt")


def generated_example_networks():
    """(name, network) of the generated solvers the tests load:
    carbon, an ion-by-ion network (mass-action descriptor); primgen, the full
    primordial network through the generator instead of the hand-written
    device code; primphoto, the primordial species with a reduced reaction
    and cooling set plus photoionisation and photoheating on redshift
    tables; oxydata, O I - O IX with the reference's Gnat & Ferland cooling and
    HM12 photoionisation/photoheating tables (data fixtures); umist, the
    hydrogen-helium reactions of the UMIST RATE12 table (its species sort after
    the energy term)."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import numpy as np
    from dengo_amd.networks import (ion_network, oxygen_data_network, primordial_network,
                                    primordial_photo_network, umist_network)
    with np.errstate(all="ignore"):
        return [("carbon", ion_network("C", seed=9)), ("primgen", primordial_network()),
                ("primphoto", primordial_photo_network()),
                ("oxydata", oxygen_data_network(ION_DATA)),
                ("umist", umist_network(RATE12_HHE))]


def build_generated_examples():
    """Writes (ChemicalNetwork.write_solver) and builds the generated solvers
    the tests load into tests/generated/<name>."""
    libs = []
    for name, net in generated_example_networks():
        out = os.path.join(GENERATED, name)
        import numpy as np
        with np.errstate(all="ignore"):
            net.write_solver(name, output_dir=out, write_h5=False)
        libs.append(build_solver(out, name))
    return libs


if __name__ == "__main__":
    print(build(verbose=True))
