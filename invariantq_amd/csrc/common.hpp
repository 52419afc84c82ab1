// Shared definitions for the gfx950 statevector engine.
//
// Amplitudes are stored array-of-structs (re, im interleaved), 16 B per
// amplitude for complex128 and 8 B for complex64, so one amplitude is one
// global_load_dwordx4 / dwordx2.  Global index bit q is qubit q (little-endian,
// the reference's convention, core/src/statevector.cpp:31-39).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace iqg {

template <typename R> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };
template <typename R> using cv = typename cvec<R>::type;

constexpr int kMaxK = 5;            // largest fused block: 5 qubits, 32x32
constexpr int kMaxN = 36;           // guards 64-bit index arithmetic
// A dispatch holds fewer than 2^32 work-items per grid dimension: a kernel
// whose work can exceed that (states above 32 qubits) loops over a grid
// capped at kMaxDispatchItems / block threads.
constexpr uint64_t kMaxDispatchItems = 0xffffffffULL;
constexpr int kRedBlocks = 1024;    // partial-sum slots for deterministic reductions
constexpr int kRedThreads = 256;

enum Precision : int { kC128 = 0, kC64 = 1 };

// Error codes surfaced through iq_last_error().
enum ErrorCode : int {
    kOk = 0,
    kErrInvalidArgument = 1,
    kErrZeroProbability = 2,
    kErrOutOfMemory = 3,
    kErrDevice = 4,
    kErrUnsupported = 5,
};

}  // namespace iqg
