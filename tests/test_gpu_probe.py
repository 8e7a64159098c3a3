"""The FP64 vector rate probe (dengo_measure_fp64_fma_rate) against MI355X's
published FP64 vector peak, 78.6 TFLOP/s (256 CUs x 128 FLOP/clock x 2.4 GHz):
a dense FMA stream reaches most of it and cannot exceed it by more than clock
variation."""
import pytest

from dengo_amd import runtime

pytestmark = pytest.mark.gpu


def test_fp64_fma_rate_near_published_peak():
    print("fp64 fma rate probe", flush=True)
    r = runtime.fp64_fma_rate(100000)
    print("measured %.1f TFLOP/s" % r, flush=True)
    assert 0.6 * 78.6 < r < 1.05 * 78.6


def test_fp64_fma_rate_rejects_bad_arguments():
    with pytest.raises(RuntimeError, match="iters"):
        runtime.fp64_fma_rate(0)
