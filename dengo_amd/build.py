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
           "dengo_server.hip"]
HEADERS = ["dengo_device.h", "be_driver.h", "be_fast.h", "ma_net.h", "dengo_solver.h", "net_oxygen.h",
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
        "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-o", LIB + ".tmp"] + srcs
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed (%d)" % r.returncode)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
