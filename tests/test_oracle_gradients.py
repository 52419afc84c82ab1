"""Pins the oracle's gradient references (FP64 parameter shift, central
differences) against closed forms and against each other, on the CPU."""

import math

import numpy as np

from oracle import oracle as orc


def test_single_ry_closed_form(oracle_lib):
    # RY(t)|0>: <Z> = cos t, d<Z>/dt = -sin t
    for t in (0.3, 1.9, 4.4):
        ops = lambda p: [([0], orc._ry(p[0]))]
        e = orc.energy(oracle_lib, 1, ops([t]), [(1.0, "Z", [0])])
        g = orc.grad_parameter_shift(oracle_lib, 1, ops, [(1.0, "Z", [0])], [t])
        assert abs(e - math.cos(t)) < 1e-14
        assert abs(g[0] + math.sin(t)) < 1e-14


def test_product_state_tfim_closed_form(oracle_lib):
    # RY(t_q) then RZ(f_q) on each qubit, no entanglers:
    # <Z_q> = cos t_q, <X_q> = sin t_q cos f_q
    n = 4
    rng = np.random.default_rng(1)
    p = rng.uniform(0, 2 * np.pi, 2 * n)
    make = lambda x: orc.hea_ops(n, 1, x, ring=False)[: 2 * n]
    terms = orc.tfim_terms(n)
    t, f = p[:n], p[n:]
    e_exact = sum(math.cos(t[i]) * math.cos(t[i + 1]) for i in range(n - 1)) + 0.5 * sum(
        math.sin(t[q]) * math.cos(f[q]) for q in range(n))
    assert abs(orc.energy(oracle_lib, n, make(p), terms) - e_exact) < 1e-13
    g_t = [-math.sin(t[q]) * ((math.cos(t[q - 1]) if q > 0 else 0) + (math.cos(t[q + 1]) if q < n - 1 else 0))
           + 0.5 * math.cos(t[q]) * math.cos(f[q]) for q in range(n)]
    g_f = [-0.5 * math.sin(t[q]) * math.sin(f[q]) for q in range(n)]
    g = orc.grad_parameter_shift(oracle_lib, n, make, terms, p)
    assert np.max(np.abs(g - np.asarray(g_t + g_f))) < 1e-13


def test_parameter_shift_matches_central_difference(oracle_lib):
    n, layers = 5, 2
    rng = np.random.default_rng(2)
    p = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    make = lambda x: orc.hea_ops(n, layers, x)
    terms = orc.tfim_terms(n)
    which = [0, 3, 7, 11, 19]
    ps = orc.grad_parameter_shift(oracle_lib, n, make, terms, p, which)
    fd = orc.grad_central_difference(oracle_lib, n, make, terms, p, h=1e-6, which=which)
    assert np.max(np.abs(ps - fd)) < 1e-8


def test_energy_matches_exact_diagonalisation_bound(oracle_lib):
    n = 6
    H = orc.tfim_dense(n)
    emin = np.linalg.eigvalsh(H)[0]
    rng = np.random.default_rng(3)
    for _ in range(3):
        p = rng.uniform(0, 2 * np.pi, 2 * n * 2)
        e = orc.energy(oracle_lib, n, orc.hea_ops(n, 2, p), orc.tfim_terms(n))
        sv = oracle_lib.statevector(n)
        for qs, m in orc.hea_ops(n, 2, p):
            (sv.apply_1q(qs[0], m) if len(qs) == 1 else sv.apply_2q(qs[0], qs[1], m))
        psi = sv.get_amplitudes()
        assert abs(e - np.vdot(psi, H @ psi).real) < 1e-12
        assert e >= emin - 1e-12
