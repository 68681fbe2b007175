"""GPU parity: the gfx950 engines through the C ABI against the CPU oracle, bit for bit.

Every case runs the same instance, colouring, ladder and seed on the GPU (drift_amd.gpu ->
libdriftpt.so) and on the oracle (oracle/pt_oracle.c) and compares spins, energies, swap
decisions, the rung permutation and the best state exactly.  Covered: integer couplings (int8,
int32, half-integer QUBO scale, fields), float couplings, both engines, several tile sizes, replica
counts that are not powers of two, R = 1 and 2, degenerate instances (one spin, no couplings at
all, an all-zero model, one pair beside isolated spins), R = 1024, zero rounds, caller colourings
(empty classes, one spin per class), the 2-D ferromagnet of config C1 (L=4 and L=16), the first
10^3 sweeps of config C2 (3-D +/-J, L=32, R=64), a batch of uneven float instances in one
persistent launch, every first-chunk width of the padded edge records (with rows longer than the
chunk), and both barrier energy paths of the bit-sliced persistent slab engine.
Reference GPU tests (tests/test_gpu.py:45-67) are reproduced last.
"""

import os

import numpy as np
import pytest

import oracle
from drift_amd import gpu
from drift_amd.builders.families import pmj_spin_glass
from drift_amd.builders.lattice import cubic_pmj_glass, random_regular_gaussian, square_ferromagnet
from drift_amd.builders.qubo import maxcut_ising, qubo_to_ising, random_graph
from drift_amd.ising import IsingModel
from drift_amd.solve import solve
from drift_amd.solvers.exact import exact_ground_state
from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not gpu.gpu_available(), reason="no gfx950 device / engine not built"),
]

THREADS = min(16, len(os.sched_getaffinity(0)))


def _announce(name):
    print(f"[gpu-parity] {name}", flush=True)


def compare(model, *, R, n_rounds, S, seed, T_min=0.1, T_max=5.0, beta=None, engine="auto",
            block=0, force_float=False, graph=True, coloring=None):
    enc = gpu.encode_problem(model, force_float=force_float)
    if coloring is None:
        k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    else:
        cptr, cspins = coloring
    if beta is None:
        beta = ladder_betas(geometric_ladder(R, T_min, T_max))
    g = gpu.parallel_tempering_gpu(
        model, n_replicas=R, n_rounds=n_rounds, sweeps_per_round=S, seed=seed, betas=beta,
        engine=engine, block_threads=block, coloring=(cptr, cspins), record_history=True,
        return_final=True, record_trace=True, force_float=force_float, graph=graph)
    o = oracle.pt_run_encoded(enc, (cptr, cspins), seed=seed, R=R, n_rounds=n_rounds,
                              sweeps_per_round=S, beta=beta, n_threads=THREADS, trace=True)
    assert g.best_E_units == o["best_units"]
    assert (g.best_round, g.best_rung) == (o["best_round"], o["best_rung"])
    assert np.array_equal(g.best_s.astype(np.int8), o["best_s"])
    assert np.array_equal(g.final_spins, o["final_spins"])
    assert np.array_equal(g.final_E, o["final_units"])
    assert np.array_equal(g.rung_of, o["rung_of"])
    assert np.array_equal(g.swap_tried, o["swap_tried"])
    assert np.array_equal(g.swap_accepted, o["swap_acc"])
    assert np.array_equal(g.round_trips, o["round_trips"])
    assert np.array_equal(g.rung_up, o["rung_up"]) and np.array_equal(g.rung_down, o["rung_down"])
    assert np.array_equal(g.best_history, o["best_history"])
    assert np.array_equal(g.rung_energy_trace, o["rung_trace"])
    # the reported energy is the energy of the reported configuration
    e = model.energy(g.best_s)
    assert g.best_E == pytest.approx(e, rel=1e-5, abs=1e-6)
    return g, o, enc


@pytest.mark.parametrize("engine", ["lds", "global", "coop"])
def test_maxcut_int8(engine):
    _announce(f"maxcut int8 {engine}")
    g, _, enc = compare(maxcut_ising(random_graph(24, p=0.5, seed=3)), R=16, n_rounds=60, S=4,
                        seed=1, T_min=0.05, T_max=4.0, engine=engine)
    assert enc.weight_type == 1 and g.engine == engine


@pytest.mark.parametrize("engine", ["lds", "global", "coop"])
def test_float_couplings_and_field(engine):
    _announce(f"float field {engine}")
    model = pmj_spin_glass(20, 7, field=0.3)
    g, _, enc = compare(model, R=20, n_rounds=80, S=3, seed=2, T_min=0.05, T_max=5.0,
                        engine=engine)
    assert not enc.is_integer


@pytest.mark.parametrize("engine", ["lds", "global", "coop"])
def test_qubo_half_integer_scale_and_int32(engine):
    _announce(f"qubo/int32 {engine}")
    rng = np.random.default_rng(4)
    Q = rng.integers(-300, 300, size=(18, 18)).astype(np.float64)
    model, _ = qubo_to_ising(Q)
    g, _, enc = compare(model, R=12, n_rounds=50, S=2, seed=9, T_min=5.0, T_max=400.0,
                        engine=engine)
    assert enc.weight_type == 2 and enc.scale in (0.5, 0.25)


@pytest.mark.parametrize("R", [1, 2, 5, 12, 33, 64, 100])
def test_replica_counts(R):
    _announce(f"R={R}")
    model = maxcut_ising(random_graph(30, p=0.3, seed=R))
    compare(model, R=R, n_rounds=30, S=2, seed=R + 100, T_min=0.2, T_max=3.0, engine="lds")
    compare(model, R=R, n_rounds=30, S=2, seed=R + 100, T_min=0.2, T_max=3.0, engine="global")
    compare(model, R=R, n_rounds=30, S=2, seed=R + 100, T_min=0.2, T_max=3.0, engine="coop")


@pytest.mark.parametrize("engine", ["lds", "global", "coop", "split"])
def test_maximum_replica_count(engine):
    """R = 1024, the most the boundary takes (2000 is refused, test_engine_errors_are_reported):
    1024-thread workgroups on the LDS, CSR and coop engines, 16 groups of 64 slots on the split
    engine."""
    _announce(f"R=1024 {engine}")
    model = maxcut_ising(random_graph(24, p=0.3, seed=6))
    g, _, _ = compare(model, R=1024, n_rounds=6, S=1, seed=77, T_min=0.1, T_max=5.0,
                      engine=engine)
    assert g.engine == engine


def _degenerate_model(kind):
    """Instances at the edge of what the boundary takes (n = 0 is refused, tests/test_lib_abi.py):
    one spin, spins with no couplings at all (nnz = 0: every row empty, the first chunk all
    padding), an all-zero model (every move has zero energy change) and one coupled pair beside
    isolated spins."""
    rng = np.random.default_rng(5)
    if kind == "one_spin_field":
        return IsingModel(np.zeros((1, 1)), np.array([2.0]))
    if kind == "no_couplings_float":
        return IsingModel(np.zeros((7, 7)), rng.normal(size=7))
    if kind == "no_couplings_int":
        return IsingModel(np.zeros((9, 9)), rng.integers(-3, 4, size=9).astype(float))
    if kind == "all_zero":
        return IsingModel(np.zeros((5, 5)), np.zeros(5))
    J = np.zeros((5, 5))   # one_pair
    J[1, 3] = J[3, 1] = -2.0
    return IsingModel(J, np.zeros(5))


@pytest.mark.parametrize("engine", ["lds", "global", "coop", "split"])
@pytest.mark.parametrize("kind", ["one_spin_field", "no_couplings_float", "no_couplings_int",
                                  "all_zero", "one_pair"])
@pytest.mark.parametrize("R", [3, 40])
def test_degenerate_instances(kind, engine, R):
    _announce(f"degenerate {kind} {engine} R={R}")
    model = _degenerate_model(kind)
    g, _, _ = compare(model, R=R, n_rounds=12, S=2, seed=31, engine=engine)
    # closed form: each uncoupled spin follows its field, the pair takes its coupling
    assert g.best_E == pytest.approx(-np.abs(model.h).sum() - np.abs(model.J).sum() / 2,
                                     rel=1e-6, abs=1e-9)


@pytest.mark.parametrize("seed", [2 ** 32, 2 ** 63, 2 ** 64 - 1])
def test_seeds_over_the_whole_64_bit_range(seed):
    """The seed keys Philox by its two 32-bit halves; seeds with the high half set or every bit
    set give the oracle's streams."""
    _announce(f"seed {seed}")
    model = maxcut_ising(random_graph(20, p=0.3, seed=2))
    for engine in ("lds", "global"):
        compare(model, R=8, n_rounds=10, S=2, seed=seed, engine=engine)


def test_int8_couplings_past_the_24_bit_position_limit():
    """A ring of 2^24 + 2 spins with +-1 couplings: int8 couplings pack a 24-bit position beside
    the weight, so an instance this large takes the two-word integer records instead; the run
    still equals the oracle's, which reads the CSR as given."""
    from drift_amd.ising import SparseIsingModel

    n = 2 ** 24 + 2
    _announce(f"ring of {n} spins, int8 couplings")
    i = np.arange(n, dtype=np.int64)
    col = np.empty(2 * n, dtype=np.int64)
    col[0::2] = (i - 1) % n
    col[1::2] = (i + 1) % n
    w_edge = np.where(np.random.default_rng(9).random(n) < 0.5, -1.0, 1.0)   # bond (i, i + 1)
    w = np.empty(2 * n)
    w[1::2] = w_edge
    w[0::2] = np.roll(w_edge, 1)
    model = SparseIsingModel(np.arange(0, 2 * n + 1, 2, dtype=np.int64), col, w, np.zeros(n),
                             check=False)
    g, _, enc = compare(model, R=2, n_rounds=2, S=1, seed=4, T_min=0.5, T_max=2.0,
                        engine="global")
    assert enc.weight_type == 1


def _colouring(model, kind):
    """Colourings a caller may pass besides the greedy one: with empty classes (at the start, in
    the middle and at the end) or one spin per class."""
    enc = gpu.encode_problem(model)
    k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    cptr = np.asarray(cptr, dtype=np.int64)
    if kind == "empty_classes":
        mid = cptr[: k // 2 + 1]
        return (np.concatenate([[0], mid, [mid[-1]], cptr[k // 2 + 1:], [cptr[-1]]]),
                np.asarray(cspins))
    return np.arange(enc.n + 1), np.random.default_rng(3).permutation(enc.n)   # one per class


@pytest.mark.parametrize("engine", ["lds", "global", "coop", "split"])
@pytest.mark.parametrize("kind", ["empty_classes", "one_per_class"])
def test_caller_colourings(kind, engine):
    _announce(f"colouring {kind} {engine}")
    model = maxcut_ising(random_graph(30, p=0.3, seed=12))
    compare(model, R=12, n_rounds=15, S=2, seed=41, engine=engine,
            coloring=_colouring(model, kind))


def _degree_profile_model(kind):
    """Graphs whose row lengths pick each first-chunk width of the padded edge records (LDS and
    split engines 2, 3, 4, and 6 for int8 couplings; word-per-position CSR engine 4, 8) and send
    long rows through the chunked remainder."""
    rng = np.random.default_rng(11)
    if kind == "lattice6_int":        # every row 6 long: width 6 (int8), no remainder
        return cubic_pmj_glass(4, 3)
    if kind == "er_int_dense":        # mean degree ~10: width 6 plus remainder chunks
        n = 90
        J = np.zeros((n, n))
        iu = np.triu_indices(n, 1)
        on = rng.random(len(iu[0])) < 0.11
        J[iu[0][on], iu[1][on]] = rng.choice([-2.0, -1.0, 1.0, 2.0], size=int(on.sum()))
        J = J + J.T
        return IsingModel(J, np.zeros(n))
    if kind == "ring_int":            # every row 2 long, two isolated spins
        n = 50
        J = np.zeros((n, n))
        for i in range(n - 2):
            j = (i + 1) % (n - 2)
            J[i, j] = J[j, i] = rng.choice([-1.0, 1.0])
        return IsingModel(J, np.zeros(n))
    if kind == "cubic_float":         # 3-regular, float couplings and field
        return random_regular_gaussian(60, seed=5)
    if kind == "er_int_field":        # mean degree ~5: width 4 plus remainder chunks
        n = 80
        J = np.zeros((n, n))
        iu = np.triu_indices(n, 1)
        on = rng.random(len(iu[0])) < 0.065
        w = rng.choice([-3.0, -2.0, -1.0, 1.0, 2.0, 3.0], size=len(iu[0]))
        J[iu[0][on], iu[1][on]] = w[on]
        J = J + J.T
        return IsingModel(J, rng.integers(-2, 3, size=n).astype(float))
    # hub_float: a chain of leaves (rows of 3) plus one hub row of n - 1
    n = 40
    J = np.zeros((n, n))
    for i in range(1, n - 1):
        J[i, i + 1] = J[i + 1, i] = rng.normal()
    J[0, 1:] = J[1:, 0] = rng.normal(size=n - 1)
    return IsingModel(J.astype(np.float32).astype(np.float64), np.zeros(n))


@pytest.mark.parametrize("engine", ["lds", "global", "split"])
@pytest.mark.parametrize("kind", ["ring_int", "cubic_float", "er_int_field", "hub_float",
                                  "lattice6_int", "er_int_dense"])
def test_first_chunk_widths(kind, engine):
    _announce(f"first chunk {kind} {engine}")
    model = _degree_profile_model(kind)
    compare(model, R=8, n_rounds=20, S=2, seed=21, engine=engine)


@pytest.mark.parametrize("block", [64, 128, 512, 1024])
def test_tile_sizes_do_not_change_results(block):
    _announce(f"block={block}")
    model = pmj_spin_glass(26, 9)
    a = compare(model, R=16, n_rounds=25, S=2, seed=3, engine="lds", block=block)[0]
    b = compare(model, R=16, n_rounds=25, S=2, seed=3, engine="global", block=block)[0]
    c = compare(model, R=16, n_rounds=25, S=2, seed=3, engine="coop", block=block)[0]
    assert np.array_equal(a.final_spins, b.final_spins)
    assert np.array_equal(a.final_spins, c.final_spins)


@pytest.mark.parametrize("case", [("int8", 20000, 32, 40, 1), ("int8", 6000, 64, 150, 3),
                                  ("f32", 4096, 16, 60, 2), ("int8", 3000, 100, 80, 1)])
def test_coop_engine_on_sparse_graphs(case):
    """The persistent CSR engine on graphs of the C3/C4 kinds: many workgroups, uneven colour
    classes, several sweeps per round and enough rounds to cycle its three energy buffers."""
    from drift_amd.builders.lattice import er_sparse_weighted, random_regular_gaussian

    kind, n, R, rounds, S = case
    _announce(f"coop {kind} n={n} R={R} rounds={rounds} S={S}")
    model = (er_sparse_weighted(n, 6.0, n) if kind == "int8"
             else random_regular_gaussian(n, n))
    g = compare(model, R=R, n_rounds=rounds, S=S, seed=n + R, T_min=0.05, T_max=5.0,
                engine="coop")[0]
    assert g.engine == "coop"


@pytest.mark.parametrize("L", [4, 16])
def test_config_c1_square_ferromagnet(L):
    _announce(f"C1 L={L}")
    beta = ladder_betas(geometric_ladder(16, 1.0, 10.0))
    compare(square_ferromagnet(L), R=16, n_rounds=400 if L == 4 else 150, S=1, seed=11,
            beta=beta, engine="lds")
    compare(square_ferromagnet(L), R=16, n_rounds=60, S=1, seed=11, beta=beta, engine="global")
    g = compare(square_ferromagnet(L), R=16, n_rounds=60, S=1, seed=11, beta=beta,
                engine="lattice")[0]
    assert g.engine == "lattice"


@pytest.mark.parametrize("case", [("lattice", 2, 6), ("lattice", 3, 4), ("lattice", 3, 6),
                                  ("lattice", 2, 32), ("slab", 2, 8), ("slab", 3, 4),
                                  ("slab", 3, 8), ("slab", 3, 16), ("slab", 2, 32),
                                  ("pslab", 2, 8), ("pslab", 3, 4), ("pslab", 3, 8),
                                  ("pslab", 3, 16), ("pslab", 2, 32)])
def test_lattice_kernel_shapes(case):
    engine, dim, L = case
    _announce(f"{engine} dim={dim} L={L}")
    model = square_ferromagnet(L) if dim == 2 else cubic_pmj_glass(L, 3)
    for R, S in ((16, 2), (64, 1), (80, 3), (5, 1)):
        g = compare(model, R=R, n_rounds=40, S=S, seed=R, T_min=0.3, T_max=4.0, engine=engine)[0]
        assert g.engine == engine


@pytest.mark.parametrize("engine", ["lds", "global", "lattice", "slab", "coop", "pslab"])
def test_zero_sweeps_per_round_only_exchanges(engine):
    _announce(f"S=0 {engine}")
    model = cubic_pmj_glass(8, 2)
    g = compare(model, R=8, n_rounds=9, S=0, seed=3, T_min=0.5, T_max=2.0, engine=engine)[0]
    assert g.best_round == 0


@pytest.mark.parametrize("engine", ["lds", "global", "lattice", "slab", "coop", "pslab", "split"])
def test_zero_rounds_report_the_initial_states(engine):
    """n_rounds = 0: the best of the initial states (round 0), with an empty history and trace,
    as the oracle's (the reference's round loop runs zero times, cuda/ising_pt.cu:213-222)."""
    _announce(f"n_rounds=0 {engine}")
    model = cubic_pmj_glass(8, 2)
    g = compare(model, R=8, n_rounds=0, S=2, seed=3, T_min=0.5, T_max=2.0, engine=engine)[0]
    assert g.best_round == 0 and g.best_history.size == 0 and g.engine == engine


def test_zero_rounds_on_the_neighbour_synchronised_slab():
    """The same on the bit-sliced slab kernel that runs without per-round records (L = 32)."""
    _announce("n_rounds=0 pslab, neighbour-synchronised")
    model = cubic_pmj_glass(32, 4)
    R = 8
    enc = gpu.encode_problem(model)
    k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    beta = ladder_betas(geometric_ladder(R, 0.5, 2.0))
    g = gpu.parallel_tempering_gpu(model, n_replicas=R, n_rounds=0, sweeps_per_round=2, seed=3,
                                   betas=beta, engine="pslab", coloring=(cptr, cspins),
                                   record_history=False, return_final=True)
    o = oracle.pt_run_encoded(enc, (cptr, cspins), seed=3, R=R, n_rounds=0, sweeps_per_round=2,
                              beta=beta, n_threads=THREADS)
    assert g.best_E_units == o["best_units"] and g.best_round == o["best_round"] == 0
    assert g.best_rung == o["best_rung"]
    assert np.array_equal(g.best_s.astype(np.int8), o["best_s"])
    assert np.array_equal(g.final_spins, o["final_spins"])
    assert np.array_equal(g.final_E, o["final_units"])
    assert np.array_equal(g.rung_of, o["rung_of"])


def test_lattice_engine_rejects_non_lattices():
    _announce("lattice rejects a random graph")
    with pytest.raises(RuntimeError, match="lattice"):
        gpu.parallel_tempering_gpu(maxcut_ising(random_graph(16, 0.5, seed=1)), n_replicas=4,
                                   n_rounds=1, engine="lattice")
    with pytest.raises(RuntimeError, match="slab"):
        gpu.parallel_tempering_gpu(square_ferromagnet(6), n_replicas=4, n_rounds=1, engine="slab")


def test_config_c2_first_thousand_sweeps():
    _announce("C2 L=32 R=64 1000 sweeps, slab engine")
    model = cubic_pmj_glass(32, 0)
    beta = ladder_betas(geometric_ladder(64, 1.0 / 3.0, 5.0))     # beta 3.0 .. 0.2
    g, o, enc = compare(model, R=64, n_rounds=1000, S=1, seed=1, beta=beta, engine="slab")
    assert enc.weight_type == 1 and g.engine == "slab"
    _announce("C2 L=32 R=64 1000 sweeps, automatic engine (persistent bit-sliced slabs)")
    g3 = compare(model, R=64, n_rounds=1000, S=1, seed=1, beta=beta, engine="auto")[0]
    assert g3.engine == "pslab"
    for engine in ("lattice", "global", "coop"):
        _announce(f"C2 L=32 R=64 100 sweeps, {engine}")
        g2 = compare(model, R=64, n_rounds=100, S=1, seed=1, beta=beta, engine=engine)[0]
        assert g2.engine == engine


def test_batch_equals_oracle_per_instance():
    _announce("batch of 48 uneven float instances")
    rng = np.random.default_rng(0)
    sizes = (2 * rng.integers(32, 129, size=48)).tolist()
    models = [random_regular_gaussian(n, seed=1000 + i) for i, n in enumerate(sizes)]
    seeds = [77 + 3 * i for i in range(len(models))]
    res = gpu.parallel_tempering_gpu_batch(models, n_replicas=16, T_min=0.1, T_max=3.0,
                                           n_rounds=40, sweeps_per_round=2, seeds=seeds,
                                           return_final=True)
    beta = ladder_betas(geometric_ladder(16, 0.1, 3.0))
    for model, seed, r in zip(models, seeds, res):
        enc = gpu.encode_problem(model)
        k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
        o = oracle.pt_run_encoded(enc, (cptr, cs), seed=seed, R=16, n_rounds=40,
                                  sweeps_per_round=2, beta=beta, n_threads=THREADS)
        assert r.best_E_units == o["best_units"]
        assert np.array_equal(r.final_spins, o["final_spins"])
        assert np.array_equal(r.rung_of, o["rung_of"])
        assert r.best_E == pytest.approx(model.energy(r.best_s), rel=1e-5)
    # one instance alone gives the same result as inside the batch
    alone = gpu.parallel_tempering_gpu(models[5], n_replicas=16, T_min=0.1, T_max=3.0,
                                       n_rounds=40, sweeps_per_round=2, seed=seeds[5],
                                       engine="lds", return_final=True)
    assert np.array_equal(alone.final_spins, res[5].final_spins)


def test_reference_gpu_reproduces_exact_on_small_maxcut():
    _announce("reference test_gpu exact")
    model = maxcut_ising(random_graph(12, p=0.5, seed=3))
    _, e_exact, _ = exact_ground_state(model)
    res = gpu.parallel_tempering_gpu(model, n_replicas=16, T_min=0.05, T_max=4.0, n_rounds=200,
                                     sweeps_per_round=4, seed=1)
    assert res.best_E == pytest.approx(e_exact)
    assert res.throughput > 0 and res.seconds > 0


def test_reference_solve_gpu_path_is_not_certified():
    _announce("reference solve gpu-pt")
    model = maxcut_ising(random_graph(20, p=0.5, seed=2))
    sol = solve(model, exact_max=16, use_gpu=True, n_replicas=16, n_rounds=80, sweeps_per_round=4,
                seed=1)
    assert sol.method == "gpu-pt" and sol.certified is False
    assert np.isclose(model.energy(sol.s), sol.energy)


def test_engine_errors_are_reported():
    _announce("error paths")
    J = np.zeros((4, 4))
    J[0, 1] = J[1, 0] = 1.0
    model = IsingModel(J, np.zeros(4))
    with pytest.raises(RuntimeError, match="colour"):
        gpu.parallel_tempering_gpu(model, n_replicas=2, n_rounds=1,
                                   coloring=(np.array([0, 4]), np.arange(4)))
    with pytest.raises(RuntimeError, match="n_replicas"):
        gpu.parallel_tempering_gpu(model, n_replicas=2000, n_rounds=1)
    # float couplings take beta below 2^100 (integer ones any finite beta)
    Jf = J.copy()
    Jf[0, 1] = Jf[1, 0] = 0.5 + 2.0 ** -20
    with pytest.raises(RuntimeError, match="2\\^100"):
        gpu.parallel_tempering_gpu(IsingModel(Jf, np.zeros(4)), n_replicas=2, n_rounds=1,
                                   betas=np.array([1.0, 2.0 ** 100]), force_float=True)
    gpu.parallel_tempering_gpu(model, n_replicas=2, n_rounds=1, betas=np.array([1.0, 2.0 ** 100]))


def test_autotune_times_candidates_and_keeps_results():
    """The autotuner times every engine and block size that can run the instance, caches the
    fastest, and a run that takes its choice still matches the oracle bit for bit."""
    from drift_amd.builders.lattice import er_sparse_weighted

    _announce("autotune")
    gpu.autotune_clear()
    model = er_sparse_weighted(6000, 6.0, 5)
    tab = gpu.autotune(model, n_replicas=32, sweeps_per_round=1, trial_rounds=20)
    ran = [c for c in tab["candidates"] if c[-1] > 0]
    engines = {c[0] for c in ran}
    assert {"global", "coop", "split"} <= engines and len(ran) >= 6
    # the default (first candidate) stays unless another is more than 3% faster
    best = min(ran, key=lambda c: c[-1])
    default = tab["candidates"][0]
    chosen = (tab["engine"], tab["block_threads"], tab["slots_per_group"], tab["couplings_in_lds"],
              tab["parts_per_group"])
    if default[-1] > 0 and best[-1] > default[-1] / 1.03:
        assert chosen == default[:5]
    else:
        assert chosen == best[:5]
    enc = gpu.encode_problem(model)
    k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    beta = ladder_betas(geometric_ladder(32, 0.1, 5.0))
    g = gpu.parallel_tempering_gpu(model, n_replicas=32, n_rounds=30, sweeps_per_round=1, seed=4,
                                   betas=beta, coloring=(cptr, cspins), autotune=True,
                                   return_final=True)
    o = oracle.pt_run_encoded(enc, (cptr, cspins), seed=4, R=32, n_rounds=30, sweeps_per_round=1,
                              beta=beta, n_threads=THREADS)
    assert g.engine == tab["engine"]
    assert g.best_E_units == o["best_units"]
    assert np.array_equal(g.final_spins, o["final_spins"])


@pytest.mark.parametrize("case", [(2, 64, 12, 2), (3, 64, 8, 2), (3, 32, 3, 3), (3, 32, 48, 1),
                                  (3, 32, 80, 1)])
def test_pslab_bit_sliced_rows(case):
    """The bit-sliced persistent slab engine on rows of one and two 32-site words (2-D and 3-D),
    with a replica count that is not a multiple of 8 and several sweeps per round; slabs of five
    or six planes (R = 48) and of eight (R = 80, round end by the whole workgroup), where the
    boundary planes are swept and published before the interior."""
    dim, L, R, S = case
    _announce(f"pslab bits dim={dim} L={L} R={R} S={S}")
    model = square_ferromagnet(L) if dim == 2 else cubic_pmj_glass(L, 7)
    g = compare(model, R=R, n_rounds=25, S=S, seed=L + R, T_min=0.3, T_max=4.0, engine="pslab")[0]
    assert g.engine == "pslab"


@pytest.mark.parametrize("case", [("int8", 32, 256), ("int8", 32, 512), ("f32", 16, 256)])
def test_global_engine_on_c3_sized_graphs(case):
    """The per-launch engine on the C3 instance itself (Erdos-Renyi N = 10^5, mean degree 6, int8
    weights, R = 32; colour classes up to 32 400 positions) and on a 3-regular N(0,1) graph of the
    same size in float32: the chain is the oracle's."""
    from drift_amd.builders.lattice import er_sparse_weighted, random_regular_gaussian

    kind, R, block = case
    _announce(f"global two-per-lane {kind} R={R} block={block}")
    model = (er_sparse_weighted(100000, 6.0, 0) if kind == "int8"
             else random_regular_gaussian(100000, 3))
    g = compare(model, R=R, n_rounds=6, S=1, seed=17, T_min=0.2, T_max=5.0, engine="global",
                block=block)[0]
    assert g.engine == "global"


@pytest.mark.parametrize("block", [64, 128, 256])
def test_pslab_workgroup_sizes(block):
    """The bit-sliced persistent slab engine with one wave per workgroup (the round end and the
    barrier in the only wave, nothing drawn ahead) and with two and four waves (the other waves
    draw the next sweep's uniforms during the barrier)."""
    _announce(f"pslab bits block={block}")
    g = compare(cubic_pmj_glass(32, 11), R=16, n_rounds=30, S=2, seed=block, T_min=0.3, T_max=4.0,
                engine="pslab", block=block)[0]
    assert g.engine == "pslab"


@pytest.mark.parametrize("path", ["atomic", "words"])
@pytest.mark.parametrize("case", [(3, 32, 16, 2), (3, 32, 64, 1), (2, 64, 12, 3)])
def test_pslab_energy_paths(case, path, monkeypatch):
    """Both ways the bit-sliced persistent engine brings a round's energy changes to the round end
    (atomic adds before the arrival flags; the changes inside the arrival and release words) give
    the oracle's chain.  The library picks one per board by timing; this fixes each in turn."""
    dim, L, R, S = case
    monkeypatch.setenv("DRIFT_PT_PSLAB_ENERGY_PATH", path)
    _announce(f"pslab energy path {path} dim={dim} L={L} R={R} S={S}")
    model = square_ferromagnet(L) if dim == 2 else cubic_pmj_glass(L, 5)
    g = compare(model, R=R, n_rounds=30, S=S, seed=R + S, T_min=0.3, T_max=4.0, engine="pslab")[0]
    assert g.engine == "pslab"


def _chain_fingerprint(r):
    return (int(r.best_E_units), int(r.best_round), r.best_s.astype(np.int8).tobytes(),
            r.final_spins.tobytes(), np.asarray(r.final_E).tobytes(), np.asarray(r.rung_of).tobytes(),
            np.asarray(r.best_history).tobytes(), np.asarray(r.swap_accepted).tobytes())


def test_pslab_energy_paths_agree_over_a_long_run(monkeypatch):
    """The two barrier energy paths on the full C2 lattice (3-D +/-J, L=32, R=64) for 2*10^4
    sweeps, far past what the oracle runs in a test: identical chains."""
    _announce("pslab energy paths, C2 lattice, 20000 sweeps")
    model = cubic_pmj_glass(32, 0)
    beta = ladder_betas(geometric_ladder(64, 1.0 / 3.0, 5.0))
    fp = {}
    for path in ("atomic", "words"):
        monkeypatch.setenv("DRIFT_PT_PSLAB_ENERGY_PATH", path)
        r = gpu.parallel_tempering_gpu(model, n_replicas=64, betas=beta, n_rounds=20000,
                                       sweeps_per_round=1, seed=5, engine="pslab",
                                       record_history=True, return_final=True)
        assert r.engine == "pslab"
        fp[path] = _chain_fingerprint(r)
    assert fp["atomic"] == fp["words"]


def test_csr_engines_agree_over_many_rounds():
    """The C3 instance (Erdos-Renyi N = 10^5, mean degree 6, int8 weights, R = 32) for 300 rounds on
    the spin-word colour step, the int8-spin colour step and the persistent CSR engine: identical
    chains."""
    from drift_amd.builders.lattice import er_sparse_weighted

    _announce("CSR engines, C3 instance, 300 rounds")
    model = er_sparse_weighted(100000, 6.0, 0)
    beta = ladder_betas(geometric_ladder(32, 0.2, 20.0))
    fps = []
    for engine, block in (("global", 256), ("global", 96), ("coop", 0)):
        r = gpu.parallel_tempering_gpu(model, n_replicas=32, betas=beta, n_rounds=300,
                                       sweeps_per_round=1, seed=9, engine=engine,
                                       block_threads=block, record_history=True, return_final=True)
        assert r.engine == engine
        fps.append(_chain_fingerprint(r))
    assert fps[0] == fps[1] == fps[2]


@pytest.mark.parametrize("engine", ["pslab", "coop"])
def test_persistent_engines_with_an_ordinary_dispatch(engine):
    """DRIFT_PT_PLAIN_LAUNCH starts the persistent engines without the cooperative-launch API;
    the chain is the same."""
    from drift_amd.builders.lattice import er_sparse_weighted

    _announce(f"plain launch {engine}")
    model = cubic_pmj_glass(32, 5) if engine == "pslab" else er_sparse_weighted(5000, 6.0, 9)
    enc = gpu.encode_problem(model)
    k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    beta = ladder_betas(geometric_ladder(16, 0.3, 4.0))
    g = gpu.parallel_tempering_gpu(model, n_replicas=16, n_rounds=20, sweeps_per_round=2, seed=6,
                                   betas=beta, engine=engine, coloring=(cptr, cspins),
                                   return_final=True, plain_launch=True)
    o = oracle.pt_run_encoded(enc, (cptr, cspins), seed=6, R=16, n_rounds=20, sweeps_per_round=2,
                              beta=beta, n_threads=THREADS)
    assert g.engine == engine
    assert g.best_E_units == o["best_units"]
    assert np.array_equal(g.best_s.astype(np.int8), o["best_s"])
    assert np.array_equal(g.final_spins, o["final_spins"])


@pytest.mark.parametrize("engine", ["global", "lattice", "slab"])
def test_direct_launches_equal_graph_replay(engine):
    """The per-launch engines replay chunks of rounds from a captured hipGraph and launch the
    remaining rounds directly; with DRIFT_PT_NO_GRAPH every round launches directly.  The round
    counts make two full replays plus a remainder.  Both runs equal the oracle and each other."""
    from drift_amd.builders.lattice import er_sparse_weighted

    _announce(f"graph replay vs direct launches, {engine}")
    model = er_sparse_weighted(3000, 6.0, 4) if engine == "global" else cubic_pmj_glass(8, 4)
    S = 2 if engine == "slab" else 1
    rounds = 1100 if engine == "slab" else (400 if engine == "global" else 700)
    fps = []
    for graph in (True, False):
        g = compare(model, R=16, n_rounds=rounds, S=S, seed=13, T_min=0.3, T_max=4.0,
                    engine=engine, graph=graph)[0]
        assert g.engine == engine
        fps.append(_chain_fingerprint(g))
    assert fps[0] == fps[1]


@pytest.mark.parametrize("case", [("pslab", "atomic"), ("pslab", "words"), ("coop", "atomic")])
def test_bounded_barrier_gives_up_from_the_host(case, monkeypatch):
    """A grid-barrier bound of 0 ms (DRIFT_PT_BARRIER_TIMEOUT_MS, set from the host) makes the
    first barrier wait that has to spin give up: the abort flag reaches every workgroup, every
    wave leaves the kernel, and the call fails with the engine's message instead of hanging.
    The next call, with the default bound, runs normally and matches the oracle."""
    from drift_amd.builders.lattice import er_sparse_weighted

    engine, path = case
    _announce(f"barrier give-up {engine} ({path})")
    monkeypatch.setenv("DRIFT_PT_PSLAB_ENERGY_PATH", path)
    model = cubic_pmj_glass(32, 5) if engine == "pslab" else er_sparse_weighted(5000, 6.0, 9)
    monkeypatch.setenv("DRIFT_PT_BARRIER_TIMEOUT_MS", "0")
    with pytest.raises(RuntimeError, match="exceeded its bound"):
        gpu.parallel_tempering_gpu(model, n_replicas=16, n_rounds=40, sweeps_per_round=2, seed=6,
                                   engine=engine)
    monkeypatch.delenv("DRIFT_PT_BARRIER_TIMEOUT_MS")
    g = compare(model, R=16, n_rounds=20, S=2, seed=6, T_min=0.3, T_max=4.0, engine=engine)[0]
    assert g.engine == engine


def test_reference_problem_file_runs_on_the_engine(tmp_path):
    """A problem.bin as drift/gpu.py writes it, solved by run_problem_bin: the result.bin reads back
    as the reference parses it, and equals a direct call with the file's parameters."""
    from drift_amd.problem_io import read_result_bin, run_problem_bin, write_problem_bin

    _announce("problem.bin")
    model = maxcut_ising(random_graph(60, p=0.2, seed=3))
    prob, res = str(tmp_path / "problem.bin"), str(tmp_path / "result.bin")
    write_problem_bin(prob, model, n_replicas=24, n_rounds=80, sweeps_per_round=3, seed=9,
                      T_min=0.1, T_max=4.0)
    g = run_problem_bin(prob, res)
    best_E, best_s = read_result_bin(res, model.n)
    assert best_E == g.best_E and np.array_equal(best_s, g.best_s)
    assert model.energy(best_s) == pytest.approx(best_E, abs=1e-9)
    d = gpu.parallel_tempering_gpu(model, n_replicas=24, T_min=np.float32(0.1), T_max=4.0,
                                   n_rounds=80, sweeps_per_round=3, seed=9)
    assert d.best_E == g.best_E
