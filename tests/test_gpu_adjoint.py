"""Adjoint gradients on the GPU against FP64 references.

References: the oracle's FP64 parameter-shift and central differences
(h = 1e-6) on the CPU at small sizes; at the full 24-qubit size, the engine's
own complex128 parameter-shift (its complex128 path is pinned to the oracle
by tests/test_gpu_parity.py).  Gradient error bar: ||g - g_ref||_2 / ||g_ref||_2 <= 1e-4
for the complex64 state with FP64 accumulation; 1e-10 absolute for complex128.
"""

import math

import numpy as np
import pytest

from invariantq_amd import gates, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def A():
    from invariantq_amd import adjoint
    from invariantq_amd.core import binding
    assert binding.HAS_NATIVE_CORE
    return adjoint


def _hea(A, n, layers, precision):
    c = A.GpuCircuit.from_ansatz(workloads.Ansatz(n, layers), precision)
    c.set_observable(workloads.tfim_terms(n))
    return c


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_hea_forward_energy_and_gradient(A, oracle_lib, precision):
    n, layers = 8, 3
    rng = np.random.default_rng(10)
    p = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    c = _hea(A, n, layers, precision)
    assert c.num_gates == layers * (3 * n)
    assert c.num_segments == layers * (n + 1)  # one pass per qubit's RY+RZ, one CZ pass
    e_ref = orc.energy(oracle_lib, n, orc.hea_ops(n, layers, p), orc.tfim_terms(n))
    e = c.expectation(p)
    psi_ref = oracle_lib.statevector(n)
    for qs, m in orc.hea_ops(n, layers, p):
        (psi_ref.apply_1q(qs[0], m) if len(qs) == 1 else psi_ref.apply_2q(qs[0], qs[1], m))
    fid = abs(np.vdot(psi_ref.get_amplitudes(), c.amplitudes())) ** 2
    e2, g = c.gradient(p)
    g_ref = orc.grad_parameter_shift(oracle_lib, n, lambda x: orc.hea_ops(n, layers, x), orc.tfim_terms(n), p)
    if precision == "complex128":
        assert abs(e - e_ref) < 1e-12 and abs(e2 - e_ref) < 1e-12
        assert fid > 1 - 1e-12
        assert np.max(np.abs(g - g_ref)) < 1e-10
    else:
        assert abs(e - e_ref) < 1e-5
        assert fid > 1 - 1e-6
        assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 1e-4


@pytest.mark.parametrize("n,wires", [(5, [0, 1, 2, 3, 4]), (17, [0, 6, 13, 16, 9])])
def test_mixed_gates_against_central_differences(A, oracle_lib, n, wires):
    """RX / U3 / fixed matrices / CX through the adjoint, vs FP64 central differences.
    At 17 qubits the tape runs as tiled stages (wires in both tile passes, two-qubit
    gates between stages)."""
    rng = np.random.default_rng(11)
    fixed1 = orc.haar_unitary(rng, 2)
    fixed2 = orc.haar_unitary(rng, 4)
    w = wires
    spec = [("rx", [w[0]], [0]), ("u3", [w[1]], [1, 2, 3]), ("cx", [w[1], w[3]], []), ("ry", [w[2]], [4]),
            ("m1", [w[4]], []), ("rz", [w[3]], [5]), ("m2", [w[4], w[0]], []), ("u3", [w[0]], [6, 7, 8]),
            ("cz", [w[2], w[4]], []), ("rx", [w[4]], [9]), ("cx", [w[4], w[2]], []), ("ry", [w[1]], [10]),
            ("rz", [w[1]], [11]), ("u3", [w[3]], [12, 13, 14])]
    print(f"mixed tape n={n}", flush=True)
    c = A.GpuCircuit(n, "complex128")
    for name, qs, ps in spec:
        if name == "m1":
            c.add(A.MATRIX1, qs, matrix=fixed1)
        elif name == "m2":
            c.add(A.MATRIX2, qs, matrix=fixed2)
        else:
            c.add(name, qs, ps)
    terms = workloads.tfim_terms(n) + [(0.7, "XY", [w[0], w[3]]), (-0.4, "YZ", [w[2], w[4]])]
    c.set_observable(terms)

    def make(p):
        ops = []
        for name, qs, ps in spec:
            if name == "rx":
                ops.append((qs, gates.rx(p[ps[0]])))
            elif name == "ry":
                ops.append((qs, gates.ry(p[ps[0]])))
            elif name == "rz":
                ops.append((qs, gates.rz(p[ps[0]])))
            elif name == "u3":
                ops.append((qs, gates.u3(p[ps[0]], p[ps[1]], p[ps[2]])))
            elif name == "cx":
                ops.append((qs, gates.CX))
            elif name == "cz":
                ops.append((qs, gates.CZ))
            elif name == "m1":
                ops.append((qs, fixed1))
            else:
                ops.append((qs, fixed2))
        return ops

    p = rng.uniform(0, 2 * np.pi, 15)
    e, g = c.gradient(p)
    assert abs(e - orc.energy(oracle_lib, n, make(p), terms)) < 1e-12
    fd = orc.grad_central_difference(oracle_lib, n, make, terms, p, h=1e-6)
    assert np.max(np.abs(g - fd)) < 1e-7
    ps = c.parameter_shift(p)
    assert np.max(np.abs(ps - orc.grad_parameter_shift(oracle_lib, n, make, terms, p))) < 1e-11


@pytest.mark.parametrize("stored", [True, False])
def test_tiled_stages_16q_against_oracle(A, oracle_lib, stored):
    """16 qubits, 3 layers (tiled stages): complex128 energy, state and every
    gradient component against the oracle's FP64 parameter shift, with the
    reverse sweep reading stored pass outputs or taking psi back."""
    n, layers = 16, 3
    rng = np.random.default_rng(16)
    p = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    c = _hea(A, n, layers, "complex128")
    if not stored:
        c.set_stored_budget(0)
    # per layer: the low qubits in one tile pass, the rest and the CZ ring in
    # a second
    assert c.num_segments == 2 * layers
    e_ref = orc.energy(oracle_lib, n, orc.hea_ops(n, layers, p), orc.tfim_terms(n))
    assert abs(c.expectation(p) - e_ref) < 1e-12
    psi_ref = oracle_lib.statevector(n)
    for qs, m in orc.hea_ops(n, layers, p):
        (psi_ref.apply_1q(qs[0], m) if len(qs) == 1 else psi_ref.apply_2q(qs[0], qs[1], m))
    assert np.max(np.abs(c.amplitudes() - psi_ref.get_amplitudes())) < 1e-12
    e, g = c.gradient(p)
    assert abs(e - e_ref) < 1e-12
    g_ref = orc.grad_parameter_shift(oracle_lib, n, lambda x: orc.hea_ops(n, layers, x), orc.tfim_terms(n), p)
    assert np.max(np.abs(g - g_ref)) < 1e-10
    c64 = _hea(A, n, layers, "complex64")
    if not stored:
        c64.set_stored_budget(0)
    e64, g64 = c64.gradient(p)
    assert abs(e64 - e_ref) < 1e-5
    assert np.linalg.norm(g64 - g_ref) / np.linalg.norm(g_ref) < 1e-4


@pytest.mark.parametrize("stored", [True, False])
def test_repeated_runs_on_one_qubit_in_a_pass(A, oracle_lib, stored):
    """U3 U3 RX on one qubit is seven sub-steps, two runs in the same tile pass
    (the second overlapping the first's register group): 17 qubits, complex128,
    against FP64 central differences and the oracle's parameter shift."""
    n = 17
    spec = []
    k = 0
    for q in (0, 3, 14, 16):
        spec += [("u3", [q], [k, k + 1, k + 2]), ("u3", [q], [k + 3, k + 4, k + 5]), ("rx", [q], [k + 6])]
        k += 7
    spec += [("cz", [0, 14], []), ("ry", [3], [k]), ("cz", [3, 16], []), ("u3", [14], [k + 1, k + 2, k + 3]),
             ("u3", [14], [k + 4, k + 5, k + 6])]
    nparams = k + 7
    c = A.GpuCircuit(n, "complex128")
    if not stored:
        c.set_stored_budget(0)
    for name, qs, ps in spec:
        c.add(name, qs, ps)
    # XX on 8 and 14 lies on no single tile pass: it goes through global gathers
    terms = workloads.tfim_terms(n) + [(0.6, "XY", [0, 14]), (-0.3, "YZ", [3, 16]), (0.45, "XX", [8, 14])]
    c.set_observable(terms)

    def make(p):
        ops = []
        for name, qs, ps in spec:
            if name == "u3":
                ops.append((qs, gates.u3(p[ps[0]], p[ps[1]], p[ps[2]])))
            elif name == "rx":
                ops.append((qs, gates.rx(p[ps[0]])))
            elif name == "ry":
                ops.append((qs, gates.ry(p[ps[0]])))
            else:
                ops.append((qs, gates.CZ))
        return ops

    p = np.random.default_rng(17).uniform(0, 2 * np.pi, nparams)
    print(f"repeated runs n={n} stored={stored}", flush=True)
    e, g = c.gradient(p)
    assert abs(e - orc.energy(oracle_lib, n, make(p), terms)) < 1e-12
    g_ref = orc.grad_parameter_shift(oracle_lib, n, make, terms, p)
    assert np.max(np.abs(g - g_ref)) < 1e-10
    fd = orc.grad_central_difference(oracle_lib, n, make, terms, p, h=1e-6)
    assert np.max(np.abs(g - fd)) < 1e-7


def test_shared_parameters_accumulate(A):
    # the same parameter on two gates: gradient is the sum of both contributions
    c = A.GpuCircuit(2, "complex128")
    c.add("ry", [0], [0])
    c.add("ry", [1], [0])
    c.set_observable([(1.0, "ZZ", [0, 1])])
    t = 0.8
    e, g = c.gradient([t])
    assert abs(e - math.cos(t) ** 2) < 1e-14
    assert abs(g[0] + 2 * math.cos(t) * math.sin(t)) < 1e-14


def test_torch_autograd_boundary(A):
    import torch
    n, layers = 6, 2
    c = _hea(A, n, layers, "complex64")
    rng = np.random.default_rng(12)
    p0 = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    for device in ("cpu", "cuda"):
        p = torch.tensor(p0, dtype=torch.float64, device=device, requires_grad=True)
        e = A.expectation(c, p)
        (2.0 * e).backward()
        e_ref, g_ref = c.gradient(p0)
        assert e.device.type == device
        assert abs(e.item() - e_ref) < 1e-12
        assert np.allclose(p.grad.cpu().numpy(), 2.0 * g_ref, atol=1e-12)


def test_full_size_24q_adjoint_vs_fp64_parameter_shift(A):
    """24 qubits, 12 layers (576 parameters, 864 gates): complex64 adjoint against
    the complex128 parameter-shift on 24 parameters spread over the layers."""
    n, layers = 24, 12
    rng = np.random.default_rng(2026)
    p = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    fast = _hea(A, n, layers, "complex64")
    exact = _hea(A, n, layers, "complex128")
    assert fast.num_gates == 864
    e64, g64 = fast.gradient(p)
    e128, g128 = exact.gradient(p)
    which = list(range(0, 576, 24))
    ps = exact.parameter_shift(p, which)
    assert np.max(np.abs(g128[which] - ps)) < 1e-10
    assert abs(e64 - e128) < 1e-4
    assert np.linalg.norm(g64 - g128) / np.linalg.norm(g128) < 1e-4


@pytest.mark.parametrize("precision,n", [("complex64", 24), ("complex128", 20), ("complex64", 17)])
def test_lambda_only_reverse_tiles_agree(A, precision, n):
    """Reverse passes over stored outputs with lambda alone in LDS (each
    thread pairs its own psi elements with lambda's partners in LDS) against
    both staged (sub-cube sums): the same overlaps summed in another order, so
    the same gradient to the precision's rounding; and timed launches cover
    every pass."""
    layers = 4
    rng = np.random.default_rng(n)
    p = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    c = _hea(A, n, layers, precision)
    print(f"lambda-only reverse tiles n={n} {precision}", flush=True)
    c.set_lambda_tiles(False)
    e0, g0 = c.gradient(p)
    c.set_lambda_tiles(True)
    c.set_timing(True)
    e1, g1 = c.gradient(p)
    kinds = [k for k, _ in c.launch_times()]
    c.set_timing(False)
    tol = 1e-12 if precision == "complex128" else 2e-5
    assert e0 == e1  # the energy comes from the forward sweep and the observable alone
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < tol
    assert kinds.count("reverse") == kinds.count("forward") > 0
    assert kinds.count("observable") == 1 and kinds.count("finalize") == 1
