"""The float Metropolis decision on the device first compares u with bounds taken from the hardware
exponential and calls the deterministic exponential (pt_expf_det, the oracle's) only inside the
band between them.  That gives the oracle's decision for every u as long as the two exponentials
differ by at most 2^-13 relative; the probe checks this for every float argument in (-87, 0]
(about 1.1e9 of them), on the device that runs the engine.  Moves that lower the energy (x >= 0)
are always accepted; the fast test accepts them outside the band as long as the hardware
exponential stays at or above 1 - 2^-13 there, which the probe checks for every float x >= 0.  The float parity tests then compare
whole runs with the oracle bit for bit.
"""

import ctypes
import math
import os

import pytest

from drift_amd import gpu

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not gpu.gpu_available(), reason="no gfx950 device / engine not built"),
]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_hardware_exp_within_the_decision_margin_for_every_argument():
    lib = ctypes.CDLL(os.path.join(ROOT, "drift_amd", "libdriftprobe.so"))
    lib.drift_probe_exp_margin.argtypes = [ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_double)]
    lib.drift_probe_exp_margin.restype = ctypes.c_int
    rel, x, low = ctypes.c_double(), ctypes.c_float(), ctypes.c_double()
    assert lib.drift_probe_exp_margin(ctypes.byref(rel), ctypes.byref(x), ctypes.byref(low)) == 0
    print(f"[exp-margin] max |p - pa| / pa = {rel.value:.3e} "
          f"(2^{math.log2(max(rel.value, 1e-300)):.2f}) at x = {x.value!r}", flush=True)
    assert rel.value <= 2.0 ** -13
    assert -87.0 < x.value <= 0.0
    print(f"[exp-margin] least hardware exp over x >= 0: {low.value!r}", flush=True)
    assert low.value >= 1.0 - 2.0 ** -13
