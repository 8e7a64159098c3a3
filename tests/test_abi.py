"""The C-ABI library loads and exports every function include/*.h declares
(no GPU work is done here)."""
import ctypes
import glob
import os
import re

from dengo_amd import runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


SINGLE = os.path.join(ROOT, "include", "dengo_single_network.h")


def declared_functions(headers=None):
    names = set()
    for h in headers or glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{[^}]*\}", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            line_start = text.rfind("\n", 0, m.start()) + 1
            line = text[line_start:m.start()]
            if line.lstrip().startswith("#") or name in ("X", "sizeof", "if", "return"):
                continue
            if re.search(r"(void|int|double|long long|char|_data|dengo_server)\s*\**\s*$", line):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    lib = runtime.load()
    single = declared_functions([SINGLE])
    decl = declared_functions() - single
    assert "dengo_evolve_primordial" in decl and "primordial_setup_data" in decl
    assert "temperature_from_mass_density_primordial" in decl
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    # the unsuffixed helpers are not exported by the multi-network library ...
    assert single == {"ensure_electron_consistency", "temperature_from_mass_density"}
    assert not any(hasattr(lib, n) for n in single)
    # ... but by every single-network solver library write_solver emits
    carbon = runtime.load_solver_library(
        os.path.join(ROOT, "tests", "generated", "carbon", "libcarbon_solver.so"), "carbon")
    assert all(hasattr(carbon, n) for n in single)


def test_bindings_cover_declarations():
    sigs = set(runtime.SIGNATURES) | set(runtime.SINGLE_NETWORK_SIGNATURES)
    assert declared_functions() <= sigs, declared_functions() - sigs


def test_no_oracle_in_product():
    # the product never loads the CPU oracle
    for path in glob.glob(os.path.join(ROOT, "dengo_amd", "**", "*.*"), recursive=True):
        if path.endswith((".py", ".hip", ".h", ".cpp")):
            text = open(path).read()
            assert "import oracle" not in text and "liboracle" not in text, path


def test_struct_layouts_match_the_headers(tmp_path):
    """Every struct the bindings mirror has the header's size and field
    offsets (gcc compiles a probe against include/*.h)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        import pytest
        pytest.skip("gcc not found")
    structs = {"dengo_options": runtime.Options, "dengo_cell_outputs": runtime.CellOutputs,
               "dengo_server_config": runtime.ServerConfig,
               "dengo_patch_result": runtime.PatchResult,
               "dengo_server_stats": runtime.ServerStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dengo_types.h"',
             '#include "dengo_server.h"', 'int main(void) {']
    for c, py in structs.items():
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (c, c))
        for f, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (c, f, c, f))
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        c, f, v = line.split()
        got[(c, f)] = int(v)
    for c, py in structs.items():
        assert got[(c, "sizeof")] == ctypes.sizeof(py), c
        for f, _ in py._fields_:
            assert got[(c, f)] == getattr(py, f).offset, (c, f)
