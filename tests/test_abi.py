"""CPU checks of the drop-in boundary (no GPU calls): the library loads, exports
every name include/invariantq.h declares, including all 49 names of the
reference core/include/c_api.h that invariantq/core/binding.py:54-299 binds at
import, and the binding's argtypes table matches the header."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "invariantq.h"
LIB = ROOT / "invariantq_amd" / "core" / "libinvariantq.so"

REFERENCE_C_API = """
iq_statevector_create iq_statevector_free iq_statevector_reset iq_statevector_apply_1q iq_statevector_apply_2q
iq_statevector_apply_fused iq_statevector_measure iq_statevector_reset_qubit iq_statevector_measure_prob
iq_statevector_get_amplitudes iq_statevector_get_probabilities iq_statevector_sample iq_statevector_entropy
iq_statevector_expval_pauli iq_density_matrix_create iq_density_matrix_free iq_density_matrix_reset
iq_density_matrix_apply_1q_unitary iq_density_matrix_apply_2q_unitary iq_density_matrix_apply_kraus_1q
iq_density_matrix_get_elements iq_density_matrix_trace iq_density_matrix_purity iq_density_matrix_entropy
iq_density_matrix_measure iq_density_matrix_reset_qubit iq_stabilizer_create iq_stabilizer_free
iq_stabilizer_reset iq_stabilizer_reset_qubit iq_stabilizer_apply_h iq_stabilizer_apply_s iq_stabilizer_apply_sdg
iq_stabilizer_apply_cnot iq_stabilizer_apply_cz iq_stabilizer_apply_swap iq_stabilizer_apply_x
iq_stabilizer_apply_y iq_stabilizer_apply_z iq_stabilizer_measure iq_stabilizer_to_string iq_mps_create
iq_mps_free iq_mps_reset iq_mps_apply_1q iq_mps_apply_2q iq_mps_entropy iq_mps_to_statevector iq_mps_measure
""".split()


def header_names():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(iq_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not LIB.exists():
        pytest.skip("libinvariantq.so not built")
    return ctypes.CDLL(str(LIB))


def test_reference_names_are_declared():
    assert len(REFERENCE_C_API) == 49
    assert set(REFERENCE_C_API) <= set(header_names())


def test_library_exports_every_declared_name(lib):
    missing = [n for n in header_names() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_matches_header():
    from invariantq_amd.core import binding
    assert set(binding.SIGNATURES) == set(header_names())
    assert binding.REFERENCE_NAMES == REFERENCE_C_API[:len(binding.REFERENCE_NAMES)] or \
        set(binding.REFERENCE_NAMES) == set(REFERENCE_C_API)
    assert binding.native_core_status()["implementation"] == "hip-gfx950"


def test_header_compiles_as_c(tmp_path):
    import subprocess
    src = tmp_path / "t.c"
    src.write_text('#include "invariantq.h"\nint main(void) { return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o", str(tmp_path / "t")],
                   check=True)
