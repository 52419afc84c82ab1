"""iq_statevector_sample_uniforms_noalloc (sampling whose scan scratch is a
static device array: no allocation, nothing waited for) against
iq_statevector_sample_uniforms on the same uniforms, through the C ABI: two
calls in a row, the second on a smaller state with fewer shots, so a stale
scratch from the first call would show; plus the checks that reject input
before any launch."""

import ctypes

import numpy as np
import pytest

from invariantq_amd import workloads

pytestmark = pytest.mark.gpu


def _fn():
    from invariantq_amd.core import binding
    f = binding.lib().iq_statevector_sample_uniforms_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = None
    return f


def _state(n, precision, seed):
    from invariantq_amd.core import binding
    sv = binding.NativeStateVector(n, precision=precision)
    for op in workloads.random_u3_cx(n, 3, seed=seed):
        if len(op.qubits) == 1:
            sv.apply_1q(op.qubits[0], op.matrix)
        else:
            sv.apply_2q(op.qubits[0], op.qubits[1], op.matrix)
    return sv


def test_noalloc_sampling_matches_allocating_entry_point():
    import torch
    from invariantq_amd.core import binding
    fn = _fn()
    for n, precision, shots, seed in [(22, "complex128", 5000, 1), (11, "complex64", 300, 2)]:
        print(f"no-alloc sampling n={n} {precision} shots={shots}", flush=True)
        sv = _state(n, precision, seed)
        u = np.random.default_rng(seed).random(shots)
        u[:3] = [0.0, np.nextafter(1.0, 0.0), 0.5]
        want = sv.sample_uniforms(u)
        du = torch.tensor(u, dtype=torch.float64, device="cuda")
        dres = torch.full((shots,), -1, dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()  # the uniforms are written on torch's stream
        fn(sv._ptr, shots, du.data_ptr(), dres.data_ptr())
        binding.check_error()
        sv.synchronize()  # the call only enqueues on the handle's stream
        got = dres.cpu().numpy().view(np.uint64)
        assert np.array_equal(got, want)
        del sv


def test_noalloc_sampling_rejects_before_launch():
    import torch
    from invariantq_amd.core import binding
    lib = binding.lib()
    fn = _fn()
    sv = _state(8, "complex128", 3)
    u = np.random.default_rng(3).random(16)
    host_out = np.zeros(16, dtype=np.uint64)
    lib.iq_clear_error()
    fn(sv._ptr, 16, u.ctypes.data, host_out.ctypes.data)  # pageable host memory: refused
    assert lib.iq_last_error() == 1
    lib.iq_clear_error()
    du = torch.tensor(u, dtype=torch.float64, device="cuda")
    fn(sv._ptr, 16, du.data_ptr(), None)
    assert lib.iq_last_error() == 1
    lib.iq_clear_error()
    big = binding.NativeStateVector(32, precision="complex64")  # 2^32 amplitudes: above the static scratch
    dres = torch.zeros(16, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    fn(big._ptr, 16, du.data_ptr(), dres.data_ptr())
    assert lib.iq_last_error() == 5
    lib.iq_clear_error()
    big.synchronize()
    assert not torch.any(dres).item()  # nothing was launched
