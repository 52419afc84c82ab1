"""Kraus sets of invariantq/noise/channels.py:11-102 as built by
invariantq_amd.gates.build_kraus_channel, pinned by trace preservation
(invariantq/verification/cptp_validator.py:55-61), closed forms of each
channel's action and the identity default for unknown names."""

import math

import numpy as np
import pytest

from invariantq_amd import fusion, gates


def apply_channel(ks, rho):
    return sum(K @ rho @ K.conj().T for K in ks)


def random_rho(rng, d):
    a = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    rho = a @ a.conj().T
    return rho / np.trace(rho)


@pytest.mark.parametrize("name,params", [("depolarizing", [0.1]), ("amp_damping", [0.3]), ("t2", [0.2]),
                                         ("bit_flip", [0.05]), ("phaseflip", [0.4]),
                                         ("thermal_relaxation", [40e-6, 60e-6, 200e-9]), ("thermal", []),
                                         ("depolarizing_2q", [0.3]), ("depol_2q", [1.5]), ("no_such_channel", [])])
def test_trace_preserving(name, params):
    ks = gates.build_kraus_channel(name, params)
    d = ks[0].shape[0]
    assert np.allclose(sum(K.conj().T @ K for K in ks), np.eye(d), atol=1e-14)


def test_unknown_name_is_identity():
    ks = gates.build_kraus_channel("Not-A-Channel", [0.7])
    assert len(ks) == 1 and np.array_equal(ks[0], np.eye(2))


def test_thermal_relaxation_closed_form():
    t1, t2, tg = 50e-6, 70e-6, 50e-9
    gamma = 1.0 - math.exp(-tg / t1)
    lam = 1.0 - math.exp(-tg * (1.0 / t2 - 1.0 / (2.0 * t1)))
    ks = gates.build_kraus_channel("thermal_relaxation", [t1, t2, tg])
    assert len(ks) == 4
    rho = np.array([[0.25, 0.3 - 0.2j], [0.3 + 0.2j, 0.75]])
    out = apply_channel(ks, rho)
    # population relaxes by gamma, coherence decays by sqrt(1-gamma) sqrt(1-lambda)
    assert out[1, 1] == pytest.approx(0.75 * (1 - gamma), abs=1e-15)
    assert out[0, 0] == pytest.approx(0.25 + 0.75 * gamma, abs=1e-15)
    assert out[0, 1] == pytest.approx((0.3 - 0.2j) * math.sqrt(1 - gamma) * math.sqrt(1 - lam), abs=1e-15)
    # defaults are T1 = 50 us, T2 = 70 us, t_gate = 50 ns
    for a, b in zip(gates.build_kraus_channel("thermal", []), ks):
        assert np.array_equal(a, b)
    # T2 is clipped to 2 T1: T2 = 500 us behaves as T2 = 100 us (no pure dephasing)
    clipped = gates.build_kraus_channel("thermal_relaxation", [t1, 500e-6, tg])
    out = apply_channel(clipped, rho)
    assert out[0, 1] == pytest.approx((0.3 - 0.2j) * math.sqrt(1 - gamma), abs=1e-15)
    # order of the products: phase-major, K_phase @ K_amp (channels.py:74-79)
    pd, ad = gates.phase_damping(lam), gates.amplitude_damping(gamma)
    assert np.allclose(ks[1], pd[0] @ ad[1]) and np.allclose(ks[2], pd[1] @ ad[0])


@pytest.mark.parametrize("p", [0.0, 0.2, 1.0])
def test_depolarizing_2q_closed_form(p):
    # sum over the 16 two-qubit Paulis: rho -> (1 - p) rho + p I / 4
    rng = np.random.default_rng(5)
    rho = random_rho(rng, 4)
    out = apply_channel(gates.build_kraus_channel("depolarizing_2q", [p]), rho)
    assert np.max(np.abs(out - ((1 - p) * rho + p * np.eye(4) / 4))) < 1e-15
    ks = gates.depolarizing_2q(p)
    assert len(ks) == 16 and ks[0].shape == (4, 4)
    # the second Pauli factor acts on the second wire: weight on I (x) X is index 1
    if p > 0:
        assert np.allclose(ks[1] / math.sqrt(p / 16), np.kron(gates.I2, gates.X))


def test_two_qubit_noise_op_fuses_as_one_channel():
    items = fusion.dm_superops_of(("noise", "depolarizing_2q", [3, 1], [0.2]))
    assert len(items) == 1 and items[0][0] == [3, 1] and items[0][1].shape == (16, 16)
    with pytest.raises(ValueError):
        fusion.dm_superops_of(("noise", "depolarizing_2q", [3], [0.2]))
