"""Gate matrices and Kraus sets fed to the engine.

Gate definitions restate invariantq/ir/qir.py:67-208 (RX/RY/RZ/U3/CX/CZ ...):
row-major, and for two-qubit gates the local index is 2*bit(q0) + bit(q1), so
CX's control is the first wire (qir.py:116-122).  Kraus sets restate
invariantq/noise/channels.py:11-102, including the identity returned for an
unknown channel name.
"""

from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np

C = np.complex128
_S2 = 1.0 / math.sqrt(2.0)

I2 = np.eye(2, dtype=C)
X = np.array([[0, 1], [1, 0]], dtype=C)
Y = np.array([[0, -1j], [1j, 0]], dtype=C)
Z = np.array([[1, 0], [0, -1]], dtype=C)
H = np.array([[_S2, _S2], [_S2, -_S2]], dtype=C)
S = np.array([[1, 0], [0, 1j]], dtype=C)
T = np.array([[1, 0], [0, np.exp(1j * math.pi / 4)]], dtype=C)

CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=C)
CZ = np.diag([1, 1, 1, -1]).astype(C)
SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=C)


def rx(theta: float) -> np.ndarray:
    """qir.py:90-94"""
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=C)


def ry(theta: float) -> np.ndarray:
    """qir.py:95-99"""
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return np.array([[c, -s], [s, c]], dtype=C)


def rz(theta: float) -> np.ndarray:
    """qir.py:100-102"""
    return np.array([[np.exp(-0.5j * theta), 0], [0, np.exp(0.5j * theta)]], dtype=C)


def u3(theta: float, phi: float, lam: float) -> np.ndarray:
    """qir.py:106-115 (Qiskit U3 form)."""
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return np.array([[c, -np.exp(1j * lam) * s],
                     [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]], dtype=C)


NAMED = {"i": I2, "id": I2, "x": X, "y": Y, "z": Z, "h": H, "s": S, "t": T,
         "cx": CX, "cnot": CX, "cz": CZ, "swap": SWAP}
PARAMETRIC = {"rx": rx, "ry": ry, "rz": rz, "u3": u3}


def gate_matrix(name: str, params: Sequence[float] = ()) -> np.ndarray:
    name = name.lower()
    if name in PARAMETRIC:
        return PARAMETRIC[name](*params)
    return NAMED[name]


# ----------------------------------------------------------------- Kraus sets

def depolarizing(p: float) -> List[np.ndarray]:
    """channels.py:18-28: {sqrt(1-3p/4) I, sqrt(p/4) X, Y, Z}."""
    p = min(1.0, max(0.0, p))
    a, b = math.sqrt(max(0.0, 1.0 - 0.75 * p)), math.sqrt(p / 4.0)
    return [a * I2, b * X, b * Y, b * Z]


def amplitude_damping(gamma: float) -> List[np.ndarray]:
    """channels.py:30-36"""
    g = min(1.0, max(0.0, gamma))
    return [np.array([[1, 0], [0, math.sqrt(1 - g)]], dtype=C),
            np.array([[0, math.sqrt(g)], [0, 0]], dtype=C)]


def phase_damping(lam: float) -> List[np.ndarray]:
    """channels.py:38-44"""
    l = min(1.0, max(0.0, lam))
    return [np.array([[1, 0], [0, math.sqrt(1 - l)]], dtype=C),
            np.array([[0, 0], [0, math.sqrt(l)]], dtype=C)]


def bit_flip(p: float) -> List[np.ndarray]:
    p = min(1.0, max(0.0, p))
    return [math.sqrt(1 - p) * I2, math.sqrt(p) * X]


def phase_flip(p: float) -> List[np.ndarray]:
    p = min(1.0, max(0.0, p))
    return [math.sqrt(1 - p) * I2, math.sqrt(p) * Z]


def thermal_relaxation(t1: float = 50e-6, t2: float = 70e-6, t_gate: float = 50e-9) -> List[np.ndarray]:
    """channels.py:61-80: T2 clipped to 2 T1; amplitude damping with
    gamma = 1 - exp(-t_gate / T1) followed by phase damping with
    lambda = 1 - exp(-t_gate * max(0, 1/T2 - 1/(2 T1))); the four products
    K_phase @ K_amp in phase-major order."""
    t2 = min(t2, 2.0 * t1)
    gamma = 1.0 - math.exp(-t_gate / t1) if t1 > 0 else 0.0
    inv_tphi = (1.0 / t2) - (1.0 / (2.0 * t1)) if (t1 > 0 and t2 > 0) else 0.0
    lam = 1.0 - math.exp(-t_gate * max(0.0, inv_tphi))
    return [kp @ ka for kp in phase_damping(lam) for ka in amplitude_damping(gamma)]


def depolarizing_2q(p: float) -> List[np.ndarray]:
    """channels.py:82-99: the 16 two-qubit Paulis P_i (x) P_j, sqrt(1 - 15p/16)
    on I (x) I and sqrt(p/16) on the rest (4x4, local index 2*bit(q0) + bit(q1))."""
    p = min(1.0, max(0.0, p))
    paulis = [I2, X, Y, Z]
    out = []
    for i in range(4):
        for j in range(4):
            w = math.sqrt(max(0.0, 1.0 - (15.0 / 16.0) * p)) if i == 0 and j == 0 else math.sqrt(p / 16.0)
            out.append(w * np.kron(paulis[i], paulis[j]))
    return out


_ONE_PARAM = {"depolarizing": depolarizing, "depol": depolarizing,
              "amplitude_damping": amplitude_damping, "amp_damping": amplitude_damping, "t1": amplitude_damping,
              "phase_damping": phase_damping, "dephasing": phase_damping, "t2": phase_damping,
              "bit_flip": bit_flip, "bitflip": bit_flip, "phase_flip": phase_flip, "phaseflip": phase_flip,
              "depolarizing_2q": depolarizing_2q, "depol_2q": depolarizing_2q}
CHANNELS = dict(_ONE_PARAM, thermal_relaxation=thermal_relaxation, thermal=thermal_relaxation)


def build_kraus_channel(name: str, params: Sequence[float]) -> List[np.ndarray]:
    """Same names, defaults and clamping as invariantq/noise/channels.py:11-102.
    Thermal relaxation reads (T1, T2, t_gate) with defaults 50 us, 70 us,
    50 ns; every other named channel reads params[0] (0 when absent); an
    unknown name gives the identity channel [I] (channels.py:101-102)."""
    key = name.lower()
    if key in ("thermal_relaxation", "thermal"):
        t1 = params[0] if len(params) > 0 else 50e-6
        t2 = params[1] if len(params) > 1 else 70e-6
        tg = params[2] if len(params) > 2 else 50e-9
        return thermal_relaxation(t1, t2, tg)
    fn = _ONE_PARAM.get(key)
    if fn is None:
        return [I2.copy()]
    return fn(params[0] if params else 0.0)
