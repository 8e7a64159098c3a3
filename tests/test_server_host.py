"""Host-side logic of the patch server, on the CPU: the arena allocator and
the admission order (dengo_amd/csrc/arena.h, compiled with g++ into a check
program), and the server's C ABI as the library exports it."""
import os
import shutil
import subprocess

import pytest

from dengo_amd import runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_arena_and_admission(tmp_path):
    exe = str(tmp_path / "arena_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-fsanitize=address,undefined",
                    "-I", os.path.join(ROOT, "dengo_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "arena_check.cc"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"


def test_server_config_defaults():
    lib = runtime.load()
    cfg = runtime.ServerConfig()
    lib.dengo_server_default_config(cfg)
    assert cfg.hbm_budget == 8 << 30 and cfg.slots == 4 and cfg.max_bypass == 8
    assert cfg.grid_share == 0 and list(cfg.precision) == [-1, -1, -1]
    assert list(cfg.grid_shares) == [0, 0, 0]
    assert cfg.dense_order == 1


def test_server_create_rejects_bad_arguments():
    lib = runtime.load()
    cfg = runtime.ServerConfig()
    lib.dengo_server_default_config(cfg)
    cfg.slots = 0
    assert not lib.dengo_server_create(cfg, None, None, None, runtime.MIXED)
    assert "slots" in runtime.last_error()
    lib.dengo_server_default_config(cfg)
    assert not lib.dengo_server_create(cfg, None, None, None, runtime.MIXED)
    assert "at least one network" in runtime.last_error()
    cfg.precision[1] = 7
    assert not lib.dengo_server_create(cfg, None, None, None, runtime.MIXED)
    assert "unknown precision 7 for network 1" in runtime.last_error()
    lib.dengo_server_default_config(cfg)
    cfg.grid_shares[2] = -1
    assert not lib.dengo_server_create(cfg, None, None, None, runtime.MIXED)
    assert "grid_shares[2]" in runtime.last_error()
