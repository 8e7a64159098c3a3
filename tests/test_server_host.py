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
    cfg.precision[1] = 7
    assert not lib.dengo_server_create(cfg, None, None, None, runtime.MIXED)
    assert "unknown precision 7 for network 1" in runtime.last_error()
    lib.dengo_server_default_config(cfg)
    cfg.grid_shares[2] = -1
    assert not lib.dengo_server_create(cfg, None, None, None, runtime.MIXED)
    assert "grid_shares[2]" in runtime.last_error()


def test_check_patch_refuses_arrays_the_server_would_overrun():
    import numpy as np
    ok = dict(state=np.zeros((5, 10)), rtol=np.zeros((5, 10)), atol=np.zeros((5, 10)))
    assert len(runtime.check_patch(10, False, **ok)) == 3
    st = np.zeros(5, dtype=np.int32)
    assert len(runtime.check_patch(10, False, status=st, **ok)) == 4
    bad = [
        dict(ok, state=np.zeros((5, 10), dtype=np.float32)),       # wrong dtype
        dict(ok, state=np.zeros((5, 11)), atol=np.zeros((5, 11))),  # wrong column count
        dict(ok, rtol=np.zeros((4, 10))),                           # rtol rows differ
        dict(ok, atol=np.zeros((10, 5)).T),                         # not contiguous
        dict(ok, state=[[0.0] * 10] * 5),                           # not an array
    ]
    for b in bad:
        with pytest.raises(ValueError):
            runtime.check_patch(10, False, **b)
    with pytest.raises(ValueError, match="int32"):
        runtime.check_patch(10, False, status=np.zeros(5, dtype=np.int64), **ok)
    with pytest.raises(ValueError, match="shape"):
        runtime.check_patch(10, False, status=np.zeros(4, dtype=np.int32), **ok)
    with pytest.raises(ValueError, match="isothermal"):
        runtime.check_patch(10, False, T=np.zeros(5), **ok)
    dense = dict(state=np.zeros((3, 128)), rtol=np.zeros((3, 128)), atol=np.zeros((3, 128)))
    with pytest.raises(ValueError, match="temperatures"):
        runtime.check_patch(128, True, **dense)
    with pytest.raises(ValueError, match="shape"):
        runtime.check_patch(128, True, T=np.zeros(4), **dense)
    assert len(runtime.check_patch(128, True, T=np.zeros(3), **dense)) == 4
