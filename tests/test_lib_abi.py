"""The C ABI on the CPU (no GPU needed): the library loads, exports every symbol include/drift_pt.h
declares with the declared struct layouts, validates problems before touching a device, reports
errors with messages, and its host-side greedy colouring equals the oracle's restatement of
drift/gpu.py:122-145."""

import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from drift_amd import _lib, gpu
from drift_amd.builders.lattice import cubic_pmj_glass, er_sparse_weighted
from drift_amd.builders.qubo import maxcut_ising, random_graph
from drift_amd.ising import IsingModel

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HEADER = os.path.join(ROOT, "include", "drift_pt.h")


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(drift_pt_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_functions()
    assert set(declared) == set(_lib.SIGNATURES), declared
    for name in declared:
        assert hasattr(lib, name), name
    version = int(re.search(r"#define DRIFT_PT_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert lib.drift_pt_abi_version() == version == _lib.ABI_VERSION
    assert lib.drift_pt_device_available() in (0, 1)
    size = re.search(r"#define DRIFT_PT_NOALLOC_BYTES \((\d+)u << (\d+)\)", open(HEADER).read())
    assert int(size.group(1)) << int(size.group(2)) == _lib.NOALLOC_BYTES


def test_ctypes_layouts_match_the_header():
    src = r"""
    #include <stdio.h>
    #include <stddef.h>
    #include "drift_pt.h"
    #define P(T, F) printf(#T "." #F " %zu\n", offsetof(T, F));
    int main(void) {
      printf("problem %zu\nparams %zu\nresult %zu\n", sizeof(drift_pt_problem),
             sizeof(drift_pt_params), sizeof(drift_pt_result));
      P(drift_pt_problem, energy_scale) P(drift_pt_problem, color_spins)
      P(drift_pt_params, beta) P(drift_pt_params, block_threads)
      P(drift_pt_result, best_round) P(drift_pt_result, rung_energy_trace)
      P(drift_pt_result, seconds) P(drift_pt_result, kernel_launches)
      return 0; }
    """
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), "-o", exe, c], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    assert int(got["problem"]) == ctypes.sizeof(_lib.Problem)
    assert int(got["params"]) == ctypes.sizeof(_lib.Params)
    assert int(got["result"]) == ctypes.sizeof(_lib.Result)
    for key, val in got.items():
        if "." in key:
            struct, field = key.split(".")
            cls = {"drift_pt_problem": _lib.Problem, "drift_pt_params": _lib.Params,
                   "drift_pt_result": _lib.Result}[struct]
            assert getattr(cls, field).offset == int(val), key


def _lib_coloring(row_ptr, col_idx):
    lib = _lib.load()
    n = row_ptr.size - 1
    cptr = np.zeros(n + 1, dtype=np.int32)
    cs = np.zeros(n, dtype=np.int32)
    k = lib.drift_pt_greedy_coloring(n, _lib.ptr(row_ptr, _lib.c_int32), _lib.ptr(col_idx, _lib.c_int32),
                                     _lib.ptr(cptr, _lib.c_int32), _lib.ptr(cs, _lib.c_int32))
    assert k >= 1
    return k, cptr[: k + 1], cs


@pytest.mark.parametrize("which", ["er", "lattice", "dense"])
def test_library_colouring_equals_oracle_and_is_valid(which):
    if which == "er":
        model = er_sparse_weighted(5000, 6.0, 3)
    elif which == "lattice":
        model = cubic_pmj_glass(8, 1)
    else:
        model = maxcut_ising(random_graph(60, p=0.5, seed=4))
    rp, ci, _ = model.to_csr()
    rp, ci = rp.astype(np.int32), ci.astype(np.int32)
    k, cptr, cs = _lib_coloring(rp, ci)
    ko, cpo, cso = oracle.greedy_coloring(rp, ci)
    assert k == ko and np.array_equal(cptr, cpo) and np.array_equal(cs, cso)
    colour = np.empty(rp.size - 1, dtype=np.int64)
    for c in range(k):
        seg = cs[cptr[c]:cptr[c + 1]]
        assert np.all(np.diff(seg) > 0)          # stable: ascending within a colour
        colour[seg] = c
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    assert not np.any(colour[rows] == colour[ci])    # no edge inside a colour
    # greedy property: each vertex takes the smallest colour unused by its lower-index neighbours
    for i in range(min(rp.size - 1, 400)):
        nb = ci[rp[i]:rp[i + 1]]
        used = set(colour[nb[nb < i]].tolist())
        assert colour[i] == min(set(range(len(used) + 1)) - used)
    if which == "lattice":   # even periodic cubic lattice: greedy gives the parity colouring
        idx = np.arange(rp.size - 1)
        parity = (idx % 8 + (idx // 8) % 8 + idx // 64) % 2
        assert k == 2 and np.array_equal(colour, parity)


def _call_run(model, R=4, coloring=None, rounds=2):
    lib = _lib.load()
    enc = gpu.encode_problem(model, coloring)
    beta = np.ones(R)
    prm = gpu._params(R, 0.1, 5.0, rounds, 1, 0, beta, "auto", 0, False, False)
    buf = gpu._Buffers(enc.n, R, rounds, False, False)
    rc = lib.drift_pt_run(enc.struct, prm, buf.struct)
    return rc, _lib.last_error(lib)


def test_validation_happens_before_the_device():
    J = np.zeros((4, 4))
    J[0, 1] = J[1, 0] = 1.0
    J[1, 2] = J[2, 1] = -2.0
    model = IsingModel(J, np.zeros(4))
    rc, msg = _call_run(model, coloring=(np.array([0, 4]), np.arange(4)))
    assert rc == _lib.ECOLOR and "colour" in msg
    rc, msg = _call_run(model, coloring=(np.array([0, 2, 4]), np.array([0, 0, 1, 2])))
    assert rc == _lib.ECOLOR and "permutation" in msg
    rc, msg = _call_run(model, R=0)
    assert rc == _lib.EINVAL and "n_replicas" in msg
    # an asymmetric CSR never reaches a device
    enc = gpu.encode_problem(model)
    enc.weight[0] = 3
    lib = _lib.load()
    buf = gpu._Buffers(4, 2, 1, False, False)
    rc = lib.drift_pt_run(enc.struct, gpu._params(2, 0.1, 5.0, 1, 1, 0, np.ones(2), "auto", 0,
                                                  False, False), buf.struct)
    assert rc == _lib.EINVAL and "symmetric" in _lib.last_error(lib)


def test_an_empty_problem_is_refused_before_the_device():
    lib = _lib.load()
    enc = gpu.encode_problem(IsingModel(np.zeros((3, 3)), np.array([1.0, -2.0, 0.0])))
    enc.struct.n = 0
    prm = gpu._params(2, 0.1, 5.0, 1, 1, 0, np.ones(2), "auto", 0, False, False)
    buf = gpu._Buffers(3, 2, 1, False, False)
    assert lib.drift_pt_run(enc.struct, prm, buf.struct) == _lib.EINVAL
    assert "n must be >= 1" in _lib.last_error(lib)


def test_a_batch_of_no_instances_is_refused_before_the_device():
    lib = _lib.load()
    for entry in (lib.drift_pt_run_batch, lib.drift_pt_run_batch_noalloc):
        assert entry(None, None, 0, None) == _lib.EINVAL
        assert "bad batch arguments" in _lib.last_error(lib)


def test_ladder_and_range_refusals_before_the_device():
    lib = _lib.load()
    J = np.zeros((4, 4))
    J[0, 1] = J[1, 0] = 1.0
    J[2, 3] = J[3, 2] = -1.0
    buf = gpu._Buffers(4, 4, 1, False, False)

    def run(model, beta=None, T=(0.1, 5.0), target=None):
        prm = gpu._params(4, T[0], T[1], 1, 1, 0, beta, "auto", 0, False, False, target=target)
        rc = lib.drift_pt_run(gpu.encode_problem(model).struct, prm, buf.struct)
        return rc, _lib.last_error(lib)

    model = IsingModel(J, np.zeros(4))
    for bad in ([1.0, 2.0, 0.0, 3.0], [1.0, -2.0, 2.0, 3.0], [1.0, np.inf, 2.0, 3.0],
                [np.nan, 1.0, 2.0, 3.0]):
        rc, msg = run(model, beta=np.array(bad))
        assert rc == _lib.EINVAL and "positive and finite" in msg
    rc, msg = run(model, T=(0.0, 5.0))
    assert rc == _lib.EINVAL and "T_min and T_max" in msg
    rc, msg = run(model, target=float("nan"))
    assert rc == _lib.EINVAL and "NaN" in msg
    # float couplings: beta below 2^100
    rc, msg = run(IsingModel(J * 0.5 + J * 2.0 ** -20, np.zeros(4)),
                  beta=np.array([1.0, 2.0, 3.0, 2.0 ** 100]))
    assert rc == _lib.ERANGE and "2^100" in msg
    # integer couplings of 2^23 with beta 1e-12: every field up to 2^23 has a nonzero threshold,
    # 4 rungs x 2^23 entries, more than the 2^24 the table takes
    rc, msg = run(IsingModel(J * 2.0 ** 23, np.zeros(4)), beta=np.full(4, 1e-12))
    assert rc == _lib.ERANGE and "threshold table" in msg


@pytest.mark.skipif(gpu.gpu_available(), reason="a gfx950 device is present")
def test_valid_problem_without_device_reports_not_built():
    J = np.zeros((4, 4))
    J[0, 1] = J[1, 0] = 1.0
    rc, msg = _call_run(IsingModel(J, np.zeros(4)))
    assert rc == _lib.ENODEV and "not built" in msg
    # no couplings at all (fields only) passes validation too
    rc, msg = _call_run(IsingModel(np.zeros((3, 3)), np.array([1.0, -2.0, 0.0])))
    assert rc == _lib.ENODEV and "not built" in msg


def test_gpu_available_is_a_bool():
    assert isinstance(gpu.gpu_available(), bool)


def test_gpu_raises_when_engine_missing():
    # reference tests/test_gpu.py:37-42
    if gpu.gpu_available():
        pytest.skip("GPU engine is present")
    J = np.zeros((4, 4))
    for i in range(3):
        J[i, i + 1] = J[i + 1, i] = 1.0
    with pytest.raises(RuntimeError, match="not built"):
        gpu.parallel_tempering_gpu(IsingModel(J, np.zeros(4)), n_replicas=2, n_rounds=2)


def test_integer_encoding():
    J = np.zeros((3, 3))
    J[0, 1] = J[1, 0] = 0.5
    J[1, 2] = J[2, 1] = -1.25
    enc = gpu.encode_problem(IsingModel(J, np.array([0.25, 0.0, 3.0])))
    assert enc.is_integer and enc.scale == 0.25
    assert np.array_equal(enc.weight * enc.scale, IsingModel(J, np.zeros(3)).to_csr()[2])
    assert np.array_equal(enc.h * enc.scale, [0.25, 0.0, 3.0])
    enc = gpu.encode_problem(IsingModel(J * 1000, np.zeros(3)))
    assert enc.weight_type == _lib.W_I32
    enc = gpu.encode_problem(IsingModel(J * 0.1, np.zeros(3)))
    assert not enc.is_integer and enc.weight.dtype == np.float32
