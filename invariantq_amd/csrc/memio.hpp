// Amplitude loads and stores with an optional non-temporal (streaming) hint.
// A pass that touches each amplitude of a large state once gains nothing from
// keeping it in L2, and caching it evicts lines still in flight; the hint is
// used only where every instruction covers whole 128-byte lines.  The
// builtins need a native vector type, so the pair is viewed as one.
#pragma once

#include "common.hpp"

namespace iqg {

template <typename R>
__device__ __forceinline__ cv<R> ld_stream(const cv<R>* p, bool nt) {
    typedef R nv __attribute__((ext_vector_type(2)));
    if (!nt) return *p;
    const nv v = __builtin_nontemporal_load(reinterpret_cast<const nv*>(p));
    cv<R> o;
    o.x = v[0];
    o.y = v[1];
    return o;
}

template <typename R>
__device__ __forceinline__ void st_stream(cv<R>* p, cv<R> v, bool nt) {
    typedef R nv __attribute__((ext_vector_type(2)));
    if (!nt) {
        *p = v;
        return;
    }
    nv w;
    w[0] = v.x;
    w[1] = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<nv*>(p));
}

// Amplitudes per thread in the whole-state read+write passes and the
// diagonal operator (a block takes spans of kPassSpan * blockDim).
constexpr int kPassSpan = 1;

// The MI355X's Infinity Cache (the memory-side cache in front of HBM).  A
// state no larger stays there between whole-state passes when they are
// cached; larger ones take the streaming (non-temporal) hint, where each
// amplitude is touched once per pass.
constexpr uint64_t kCacheStateMax = 256ULL << 20;
inline bool beyond_cache(uint64_t count, uint64_t elem_bytes) { return count * elem_bytes > kCacheStateMax; }

}  // namespace iqg
