"""Batched persistent launch against the single-circuit adjoint engine and the
oracle: both code paths (state resident in LDS, and two tiled HBM passes per
layer), both precisions, determinism, and the autotuner."""

import numpy as np
import pytest

from invariantq_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    from invariantq_amd import adjoint, batch
    return adjoint, batch


def _single(adjoint, n, layers, params, precision="complex128"):
    c = adjoint.GpuCircuit.from_ansatz(workloads.Ansatz(n, layers), precision)
    c.set_observable(workloads.tfim_terms(n))
    return c.gradient(params)


SHAPES = [(2, 3), (3, 2), (5, 4), (8, 3), (12, 2), (14, 2), (15, 3), (16, 1), (13, 2)]
TILES = {8: 6, 14: 8, 15: 8, 16: 9, 12: 12, 13: 7}  # forces both paths: n <= width resident, else tiled


def _specs(seed):
    rng = np.random.default_rng(seed)
    return [(n, l, rng.uniform(0, 2 * np.pi, 2 * n * l)) for n, l in SHAPES]


def test_batch_matches_single_circuit_c128(M):
    adjoint, batch = M
    specs = _specs(1)
    r = batch.BatchRunner("complex128")
    e, g, ms = r.run(specs, tile_bits=TILES, block_threads=256, solo=False)
    assert ms > 0
    for (n, l, p), ei, gi in zip(specs, e, g):
        e_ref, g_ref = _single(adjoint, n, l, p)
        assert abs(ei - e_ref) < 1e-10, (n, l)
        assert np.max(np.abs(gi - g_ref)) < 1e-10, (n, l)


def test_batch_small_against_oracle(M, oracle_lib):
    _, batch = M
    n, l = 5, 3
    rng = np.random.default_rng(4)
    p = rng.uniform(0, 2 * np.pi, 2 * n * l)
    e, g, _ = batch.BatchRunner("complex128").run([(n, l, p)], solo=False)
    terms = orc.tfim_terms(n)
    assert abs(e[0] - orc.energy(oracle_lib, n, orc.hea_ops(n, l, p), terms)) < 1e-12
    ps = orc.grad_parameter_shift(oracle_lib, n, lambda x: orc.hea_ops(n, l, x), terms, p)
    assert np.max(np.abs(g[0] - ps)) < 1e-11


def test_batch_complex64_error_and_determinism(M):
    adjoint, batch = M
    specs = _specs(2)
    r = batch.BatchRunner("complex64")
    e1, g1, _ = r.run(specs, tile_bits=TILES, block_threads=512, solo=False)
    e2, g2, _ = r.run(list(reversed(specs)), tile_bits=TILES, block_threads=512, solo=False)
    for (n, l, p), a, b, ga, gb in zip(specs, e1, reversed(e2), g1, reversed(g2)):
        assert a == b and np.array_equal(ga, gb)  # independent of queue order
        e_ref, g_ref = _single(adjoint, n, l, p)
        assert abs(a - e_ref) < 1e-4
        assert np.linalg.norm(ga - g_ref) / np.linalg.norm(g_ref) < 1e-4


def test_batch_validation(M):
    _, batch = M
    r = batch.BatchRunner("complex64")
    with pytest.raises(ValueError):
        r.run([(20, 2, np.zeros(80))], tile_bits={20: 9}, solo=False)  # 2 * width < n
    with pytest.raises(ValueError):
        r.run([(4, 2, np.zeros(3))], solo=False)


def test_autotuner_table(M, tmp_path):
    _, batch = M
    rng = np.random.default_rng(5)
    specs = [(n, 2, rng.uniform(0, 6.3, 4 * n)) for n in (10, 14) for _ in range(16)]
    r = batch.BatchRunner("complex64")
    t = batch.autotune(r, specs, cache=tmp_path / "tune.json", per_n=16)
    assert set(t["tile_bits"]) == {10, 14} and t["block_threads"] in (256, 512)
    assert t["group_width"] in (2, 3, 4) and set(t["timings"]["group_width"]) == {"2", "3", "4"}
    again = batch.autotune(r, specs, cache=tmp_path / "tune.json", per_n=16)
    assert again["cached"] and again["tile_bits"] == t["tile_bits"] and again["group_width"] == t["group_width"]
    e, g, _ = r.run(specs, tile_bits=t["tile_bits"], block_threads=t["block_threads"], group_width=t["group_width"],
                    solo=False)
    assert np.all(np.isfinite(e))


@pytest.mark.parametrize("precision,widths", [("complex64", (2, 3, 4)), ("complex128", (2, 3))])
def test_group_widths_agree(M, precision, widths):
    """Every fusion width the autotuner may pick gives the same energies and
    gradients (the rotations of a register group commute; only the order of
    the sums changes), on resident and tiled circuits."""
    _, batch = M
    specs = _specs(7)
    r = batch.BatchRunner(precision)
    ref_e, ref_g, _ = r.run(specs, tile_bits=TILES, block_threads=512, solo=False)
    tol = 1e-12 if precision == "complex128" else 2e-5
    for w in widths:
        e, g, _ = r.run(specs, tile_bits=TILES, block_threads=512, group_width=w, solo=False)
        for (n, l, _), a, b, ga, gb in zip(specs, e, ref_e, g, ref_g):
            assert abs(a - b) < tol, (w, n, l)
            assert np.max(np.abs(ga - gb)) < tol * max(1.0, np.max(np.abs(gb))), (w, n, l)
    with pytest.raises(ValueError):
        r.run(specs[:1], group_width=5, solo=False)


@pytest.mark.parametrize("precision,tiles,threads", [("complex64", 12, 256), ("complex64", 13, 512),
                                                     ("complex128", 11, 256), ("complex128", 12, 512)])
def test_register_group_launches(M, precision, tiles, threads):
    """Tiles of exactly (elements per thread) x (threads) amplitudes with the
    widest fusion width: each thread's loaded elements form the sub-cube of the
    highest tile bits, whose rotations (forward) and reverse steps run in
    registers at load.  Circuits with 13-20 qubits (tiled, both passes on the
    register group, and n = 17 where pass B has no other group) against the
    single-circuit complex128 adjoint."""
    adjoint, batch = M
    rng = np.random.default_rng(tiles * 1000 + threads)
    shapes = [(13, 3), (15, 2), (17, 2), (18, 1), (20, 1)]
    specs = [(n, l, rng.uniform(0, 2 * np.pi, 2 * n * l)) for n, l in shapes if n <= 2 * tiles]
    r = batch.BatchRunner(precision)
    e, g, _ = r.run(specs, tile_bits={n: tiles for n, _, _ in specs}, block_threads=threads, solo=False)
    tol = 1e-10 if precision == "complex128" else 1e-4
    for (n, l, p), ei, gi in zip(specs, e, g):
        e_ref, g_ref = _single(adjoint, n, l, p)
        assert abs(ei - e_ref) < tol * max(1.0, abs(e_ref)), (n, l)
        assert np.linalg.norm(gi - g_ref) / np.linalg.norm(g_ref) < tol, (n, l)


def test_batch_1024_circuits_every_size_stratified(M):
    """A launch of 1024 circuits with n in 10..20 and up to 16 layers (the
    batched workload's distribution), autotuned; two circuits per qubit count
    (the deepest and one at random) against the single-circuit complex128
    adjoint at 1e-4 relative."""
    adjoint, batch = M
    specs = workloads.batch_specs(1024, seed=11)
    ns = sorted({s[0] for s in specs})
    assert ns == list(range(10, 21)) and max(s[1] for s in specs) == 16
    r = batch.BatchRunner("complex64")
    tune = batch.autotune(r, specs, cache=None, per_n=16)
    print(f"1024-circuit batch, tiles {tune['tile_bits']}, {tune['block_threads']} threads, "
          f"fusion width {tune['group_width']}", flush=True)
    e, g, ms = r.run(specs, tile_bits=tune["tile_bits"], block_threads=tune["block_threads"],
                     group_width=tune["group_width"], solo=False)
    assert ms > 0 and np.all(np.isfinite(e))
    rng = np.random.default_rng(12)
    for n in ns:
        idx = [i for i, s in enumerate(specs) if s[0] == n]
        deepest = max(idx, key=lambda i: specs[i][1])
        other = int(rng.choice([i for i in idx if i != deepest]))
        for i in (deepest, other):
            _, l, p = specs[i]
            e_ref, g_ref = _single(adjoint, n, l, p)
            assert abs(e[i] - e_ref) < 1e-4 * max(1.0, abs(e_ref)), (n, l)
            assert np.linalg.norm(g[i] - g_ref) / np.linalg.norm(g_ref) < 1e-4, (n, l)


def test_few_large_circuits_take_the_single_circuit_engine(M):
    """A batch whose large size classes hold a few circuits each: those go
    through the single-circuit engine (a circuit of the persistent launch
    holds one compute unit), the small ones through the launch.  Every
    energy and gradient against the single-circuit complex128 adjoint, and
    the launch's share bit for bit as a launch of those circuits alone."""
    adjoint, batch = M
    rng = np.random.default_rng(31)
    specs = [(20, 5, rng.uniform(0, 2 * np.pi, 200)), (19, 10, rng.uniform(0, 2 * np.pi, 380))]
    specs += [(10, 4, rng.uniform(0, 2 * np.pi, 80)) for _ in range(20)]
    specs += [(20, 2, rng.uniform(0, 2 * np.pi, 80)), (19, 10, rng.uniform(0, 2 * np.pi, 380))]
    r = batch.BatchRunner("complex64")
    assert r.solo_classes(specs) == [19, 20]
    # one shallow 18-qubit circuit stays in the launch: making its circuit
    # costs more than the launch would
    assert r.solo_classes([(18, 3, rng.uniform(0, 2 * np.pi, 108))]) == []
    e, g, ms = r.run(specs)
    assert r.last_solo == 4 and ms > 0
    small = [s for s in specs if s[0] == 10]
    e_b, g_b, _ = r.run(small, solo=False)
    k = 0
    for i, (n, l, p) in enumerate(specs):
        e_ref, g_ref = _single(adjoint, n, l, p)
        assert abs(e[i] - e_ref) < 1e-5
        assert np.linalg.norm(g[i] - g_ref) <= 1e-4 * max(1.0, np.linalg.norm(g_ref))
        if n == 10:
            assert e[i] == e_b[k] and np.array_equal(g[i], g_b[k])
            k += 1
    with pytest.raises(ValueError):
        r.run([(20, 2, np.zeros(79))])
