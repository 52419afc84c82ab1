// Stabilizer-tableau and MPS entry points of reference core/include/c_api.h:40-65.
//
// The reference binding binds every one of the 49 C names at import time
// (invariantq/core/binding.py:54-299), so the drop-in library must export them.
// These two engines are not dense statevector work and are not implemented on
// the GPU: *_create returns NULL (the reference binding turns that into
// MemoryError, binding.py:483-486 and :544-547) and sets iq_last_error() to
// "unsupported"; every other call is a no-op returning 0.

#include <cstddef>
#include <cstdint>
#include <cstring>

#include "common.hpp"

namespace iqg {
void set_error(int code, const char* fmt, ...);
}

extern "C" {

void* iq_stabilizer_create(size_t) {
    iqg::set_error(iqg::kErrUnsupported, "stabilizer tableau is not provided by the hip-gfx950 backend");
    return nullptr;
}
void iq_stabilizer_free(void*) {}
void iq_stabilizer_reset(void*) {}
void iq_stabilizer_reset_qubit(void*, size_t, double) {}
void iq_stabilizer_apply_h(void*, size_t) {}
void iq_stabilizer_apply_s(void*, size_t) {}
void iq_stabilizer_apply_sdg(void*, size_t) {}
void iq_stabilizer_apply_cnot(void*, size_t, size_t) {}
void iq_stabilizer_apply_cz(void*, size_t, size_t) {}
void iq_stabilizer_apply_swap(void*, size_t, size_t) {}
void iq_stabilizer_apply_x(void*, size_t) {}
void iq_stabilizer_apply_y(void*, size_t) {}
void iq_stabilizer_apply_z(void*, size_t) {}
int iq_stabilizer_measure(void*, size_t, double) { return 0; }
void iq_stabilizer_to_string(void*, char* buffer, size_t max_len) {
    if (buffer && max_len) buffer[0] = '\0';
}

void* iq_mps_create(size_t, size_t, double) {
    iqg::set_error(iqg::kErrUnsupported, "MPS is not provided by the hip-gfx950 backend");
    return nullptr;
}
void iq_mps_free(void*) {}
void iq_mps_reset(void*) {}
void iq_mps_apply_1q(void*, size_t, const double*, const double*) {}
void iq_mps_apply_2q(void*, size_t, size_t, const double*, const double*) {}
double iq_mps_entropy(void*, size_t) { return 0.0; }
void iq_mps_to_statevector(void*, double*, double*) {}
int iq_mps_measure(void*, size_t, double) { return 0; }

}  // extern "C"
