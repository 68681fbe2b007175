"""Host logic of bench.py that needs no GPU: the refusal of several GPUs, and the algorithmic
bytes per attempt the roofline figures are computed from (the per-attempt formula of the
measurement plan: own spin read and write, d neighbour spins, d edge records, two row pointers,
the field; lattices read one bond-sign byte instead of edge records)."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from drift_amd import gpu  # noqa: E402
from drift_amd.builders.lattice import cubic_pmj_glass, er_sparse_weighted  # noqa: E402


def test_several_gpus_are_refused_with_a_message():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 2
    assert "one GPU only" in out.stderr
    assert out.stdout == ""


def test_lattice_bytes_per_attempt():
    # 3-D lattice, d = 6 int8 spins: 7 reads + 1 write + 1 sign byte = 9 B
    model = cubic_pmj_glass(8, 0)
    enc = gpu.encode_problem(model)
    assert bench.algorithmic_bytes_per_attempt(model, enc, "pslab", 64) == pytest.approx(9.0)
    assert bench.lattice_side(model.n) == 8


def test_csr_bytes_per_attempt_by_spin_format():
    model = er_sparse_weighted(4000, 6.0, 0)
    enc = gpu.encode_problem(model)
    d = model.nnz / model.n
    # int8 weights pack with the column into 4-byte records; no field
    assert enc.weight.dtype.itemsize == 1 and enc.h is None
    int8_spins = bench.algorithmic_bytes_per_attempt(model, enc, "global", 64)
    words = bench.algorithmic_bytes_per_attempt(model, enc, "global", 32)
    assert int8_spins == pytest.approx((d + 2) + 4 * d + 8)
    # one bit of a 32-bit word per spin when R <= 32
    assert words == pytest.approx((d + 2) / 8 + 4 * d + 8)
    assert "one 32-bit word per position" in bench.engine_format("global", model, enc, 32)
    assert "int8 spins [position][replica]" in bench.engine_format("global", model, enc, 64)
