// C-ABI of libinvariantq.so on gfx950: statevector and density-matrix handles
// whose amplitudes live in HBM.
//
// Boundary contract (reference core/include/c_api.h:10-38, core/src/api.cpp:14-192,
// invariantq/core/binding.py:302-476):
//   * same names and signatures; matrices arrive split re/im, row-major, and are
//     copied before the call returns (the caller keeps ownership);
//   * gate applications are enqueued on the handle's stream and return at once;
//     every call that hands data back (amplitudes, probabilities, samples,
//     scalars) waits for the stream first, so the observable behaviour is the
//     reference's synchronous one;
//   * errors never cross extern "C" as exceptions: invalid qubits, k > 5 and
//     zero-probability collapses leave the state unchanged and set a
//     thread-local code read by iq_last_error() (the reference throws there,
//     core/src/statevector.cpp:27-29,50-52,100-102,173-175, which aborts the process).
//
// A density matrix is the 2n-qubit vector of row-major rho (element (r, c) at
// r * 2^n + c, reference core/src/density_matrix.cpp:9): row qubit q is bit n+q,
// column qubit q is bit q.  U rho U^dagger is the superoperator U (x) conj(U) on
// bits (n+q...), (q...) and a Kraus set is sum_k K (x) conj(K) — one HBM pass
// each instead of the reference's K+2 full copies.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "jacobi_rows.hpp"
#include "kernels.hpp"

using namespace iqg;

namespace iqg {

static thread_local int t_err = kOk;
static thread_local char t_msg[256] = "";
static int g_gate_path = kPathAuto;

void set_error(int code, const char* fmt, ...) {
    t_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_msg, sizeof(t_msg), fmt, ap);
    va_end(ap);
}

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error(kErrDevice, "%s: %s", what, hipGetErrorString(e));
    return false;
}

int gate_path_override() { return g_gate_path; }

DeviceGuard::DeviceGuard(int dev) {
    hipGetDevice(&prev_);
    if (prev_ != dev) hipSetDevice(dev);
    dev_ = dev;
}
DeviceGuard::~DeviceGuard() {
    if (prev_ != dev_) hipSetDevice(prev_);
}

// ----------------------------------------------------------------- Engine

bool Engine::init(int nqubits, int precision) {
    n = nqubits;
    prec = precision;
    dim = 1ULL << n;
    elem = prec == kC128 ? 16 : 8;
    hipGetDevice(&dev);
    if (!hip_ok(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream create")) return false;
    if (hipMalloc(&amp, dim * elem) != hipSuccess) {
        set_error(kErrOutOfMemory, "cannot allocate %llu bytes of device memory", (unsigned long long)(dim * elem));
        amp = nullptr;
        return false;
    }
    const size_t red = (size_t)kRedBlocks * 64 * sizeof(double);
    if (!hip_ok(hipMalloc(&red_part, red), "alloc") || !hip_ok(hipMalloc(&red_out, 128 * sizeof(double)), "alloc") ||
        !hip_ok(hipHostMalloc((void**)&host_out, 128 * sizeof(double)), "pinned alloc"))
        return false;
    slot_bytes = std::max(gate_matrix_bytes(kC128, kMaxK, true), mfma_fragment_bytes(kC128));
    slots.resize(kSlots);
    for (auto& s : slots) {
        // coherent: the upload kernel reads the slot uncached, so a reused slot
        // is never served from a stale GPU cache line
        if (!hip_ok(hipHostMalloc(&s.host, slot_bytes, hipHostMallocCoherent), "pinned alloc") ||
            !hip_ok(hipHostGetDevicePointer(&s.host_dev, s.host, 0), "pinned map") ||
            !hip_ok(hipMalloc(&s.dev, slot_bytes), "alloc") ||
            !hip_ok(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming), "event"))
            return false;
    }
    return reset_basis(0);
}

Engine::~Engine() {
    if (stream) hipStreamSynchronize(stream);
    if (seq.host) hipHostFree(seq.host);
    if (seq.dev) hipFree(seq.dev);
    if (seq.ev) hipEventDestroy(seq.ev);
    for (auto& s : slots) {
        if (s.host) hipHostFree(s.host);
        if (s.dev) hipFree(s.dev);
        if (s.ev) hipEventDestroy(s.ev);
    }
    if (amp) hipFree(amp);
    if (red_part) hipFree(red_part);
    if (red_out) hipFree(red_out);
    if (host_out) hipHostFree(host_out);
    if (work) hipFree(work);
    if (stage) hipFree(stage);
    if (pinned) hipHostFree(pinned);
    for (hipEvent_t ev : pinned_ev)
        if (ev) hipEventDestroy(ev);
    if (stream) hipStreamDestroy(stream);
}

bool Engine::reset_basis(uint64_t index) {
    return hip_ok(launch_init_basis(amp, n, prec, index, stream), "init");
}

const void* Engine::upload_matrix(int K, bool m3, const double* re, const double* im, const uint32_t* mfma_q) {
    Slot& s = slots[next_slot];
    next_slot = (next_slot + 1) % kSlots;
    if (s.pending) hipEventSynchronize(s.ev);
    size_t bytes;
    if (mfma_q) {
        pack_mfma_fragments(prec, n, mfma_q, re, im, s.host);
        bytes = mfma_fragment_bytes(prec);
    } else {
        pack_gate_matrix(prec, K, m3, re, im, s.host);
        bytes = gate_matrix_bytes(prec, K, m3);
    }
    // a copy kernel on this stream, not hipMemcpyAsync: the copy engine adds
    // tens of microseconds of cross-queue signalling before the gate kernel
    if (!hip_ok(launch_upload(s.dev, s.host_dev, (bytes + 15) & ~size_t(15), stream), "matrix upload")) return nullptr;
    hip_ok(hipEventRecord(s.ev, stream), "event record");
    s.pending = true;
    return s.dev;
}

static bool is_diagonal(int K, const double* re, const double* im) {
    const int d = 1 << K;
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c)
            if (r != c && (re[r * d + c] != 0.0 || im[r * d + c] != 0.0)) return false;
    return true;
}

// A K-qubit matrix (MSB-first local index: bit K-1-m of l is b(q[m])) of the
// form blockdiag(I, U), U acting on one qubit where all the others (the
// controls) are 1: returns that target's list position and U, or -1.
static int controlled_form(int K, const double* re, const double* im, double* ure, double* uim) {
    const int d = 1 << K;
    for (int tpos = 0; tpos < K; ++tpos) {
        const int tb = 1 << (K - 1 - tpos), cb = (d - 1) & ~tb;  // local bits of target and controls
        bool ok = true;
        for (int r = 0; r < d && ok; ++r)
            for (int c = 0; c < d && ok; ++c)
                if ((r & cb) != cb || (c & cb) != cb)
                    ok = re[r * d + c] == (r == c ? 1.0 : 0.0) && im[r * d + c] == 0.0;
        if (!ok) continue;
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) {
                const int R = cb | (r ? tb : 0), C = cb | (c ? tb : 0);
                ure[r * 2 + c] = re[R * d + C];
                uim[r * 2 + c] = im[R * d + C];
            }
        return tpos;
    }
    return -1;
}

bool Engine::check_gate(int K, const uint32_t* q) const {
    if (K < 1 || K > kMaxK || K > n) {
        set_error(kErrInvalidArgument, "gate on %d qubits is not supported (1..5, <= n)", K);
        return false;
    }
    for (int m = 0; m < K; ++m) {
        if ((int)q[m] >= n) {
            set_error(kErrInvalidArgument, "qubit index %u out of range for %d qubits", q[m], n);
            return false;
        }
        for (int j = 0; j < m; ++j)
            if (q[j] == q[m]) {
                set_error(kErrInvalidArgument, "repeated qubit %u in gate", q[m]);
                return false;
            }
    }
    return true;
}

// Kernel choice for one (already checked) gate: diagonal pass, matrix by
// value (K <= 3 direct), or a device copy of `bytes` (MFMA fragments or the
// packed row layout).
Engine::GatePlan Engine::plan_gate(int K, const uint32_t* q, const double* re, const double* im, int path) const {
    GatePlan p;
    p.g.K = K;
    for (int m = 0; m < K; ++m) p.g.q[m] = q[m];
    p.g.re = re;
    p.g.im = im;
    p.g.path = path;
    if (path == kPathAuto && is_diagonal(K, re, im)) {
        p.kind = GatePlan::kDiag;
        return p;
    }
    // controlled one-qubit gates (one or two controls): half or a quarter of
    // the state, once every control bit separates whole 256-byte spans (below
    // that the skipped amplitudes share cache lines with the rest)
    if (path == kPathAuto && (K == 2 || K == 3)) {
        const int tpos = controlled_form(K, re, im, p.ure, p.uim);
        bool high = tpos >= 0;
        for (int m = 0; m < K && high; ++m)
            if (m != tpos) {
                high = q[m] >= 4;
                p.ctrl[p.nctrl++] = q[m];
            }
        if (high) {
            p.kind = GatePlan::kControlled;
            p.targ = q[tpos];
            return p;
        }
        p.nctrl = 0;
    }
    int chosen = choose_path(n, prec, p.g);
    if (chosen == kPathMfma && (K != 5 || n < 9)) chosen = kPathDirect;
    p.g.path = chosen;
    if (chosen == kPathMfma) {
        p.kind = GatePlan::kDevice;
        p.mfma = true;
        p.bytes = mfma_fragment_bytes(prec);
    } else if (K >= 4 || chosen == kPathTiled) {
        p.kind = GatePlan::kDevice;
        p.m3 = use_three_mult(K);
        p.bytes = gate_matrix_bytes(prec, K, p.m3);
    } else {
        p.kind = GatePlan::kValue;
    }
    return p;
}

void Engine::pack_plan(const GatePlan& p, void* host_dst) const {
    if (p.mfma) pack_mfma_fragments(prec, n, p.g.q, p.g.re, p.g.im, host_dst);
    else pack_gate_matrix(prec, p.g.K, p.m3, p.g.re, p.g.im, host_dst);
}

bool Engine::launch_plan(GatePlan& p, const void* dmat) {
    if (p.kind == GatePlan::kDiag) {
        const int d = 1 << p.g.K;
        double dre[32], dimv[32];
        for (int l = 0; l < d; ++l) {
            dre[l] = p.g.re[l * d + l];
            dimv[l] = p.g.im[l * d + l];
        }
        return hip_ok(launch_diag(amp, n, prec, stream, p.g.K, p.g.q, dre, dimv), "diagonal gate launch");
    }
    if (p.kind == GatePlan::kControlled)
        return hip_ok(launch_controlled(amp, n, prec, stream, p.ctrl, p.nctrl, p.targ, p.ure, p.uim),
                      "controlled gate launch");
    p.g.dmat = dmat;
    return hip_ok(launch_gate(amp, n, prec, stream, p.g), "gate launch");
}

bool Engine::apply(int K, const uint32_t* q, const double* re, const double* im, int path) {
    if (!check_gate(K, q)) return false;
    GatePlan p = plan_gate(K, q, re, im, path);
    const void* dmat = nullptr;
    if (p.kind == GatePlan::kDevice) {
        dmat = upload_matrix(p.g.K, p.m3, p.g.re, p.g.im, p.mfma ? p.g.q : nullptr);
        if (!dmat) return false;
    }
    return launch_plan(p, dmat);
}

// A whole gate sequence: every gate checked first (an invalid one leaves the
// state untouched), every device matrix packed into one pinned buffer and
// uploaded once, then the gate kernels back to back on the stream.
bool Engine::apply_sequence(size_t count, const int* ks, const uint32_t* qubits, const double* re, const double* im,
                            int path) {
    std::vector<GatePlan> plans(count);
    std::vector<size_t> offs(count, 0);
    size_t qo = 0, mo = 0, bytes = 0;
    for (size_t i = 0; i < count; ++i) {
        const int K = ks[i];
        if (K < 1 || K > kMaxK) {
            set_error(kErrInvalidArgument, "gate %zu: %d qubits is not supported (1..5)", i, K);
            return false;
        }
        if (!check_gate(K, qubits + qo)) return false;
        plans[i] = plan_gate(K, qubits + qo, re + mo, im + mo, path);
        if (plans[i].kind == GatePlan::kDevice) {
            offs[i] = bytes;
            bytes += (plans[i].bytes + 255) & ~size_t(255);
        }
        qo += (size_t)K;
        mo += (size_t)1 << (2 * K);
    }
    if (bytes) {
        if (seq.pending) hip_ok(hipEventSynchronize(seq.ev), "sequence wait");
        seq.pending = false;
        if (bytes > seq.cap) {
            if (seq.host) hipHostFree(seq.host);
            if (seq.dev) hipFree(seq.dev);
            seq.host = seq.host_dev = seq.dev = nullptr;
            seq.cap = 0;
            if (!hip_ok(hipHostMalloc(&seq.host, bytes, hipHostMallocCoherent), "pinned alloc") ||
                !hip_ok(hipHostGetDevicePointer(&seq.host_dev, seq.host, 0), "pinned map") ||
                !hip_ok(hipMalloc(&seq.dev, bytes), "alloc"))
                return false;
            seq.cap = bytes;
        }
        if (!seq.ev && !hip_ok(hipEventCreateWithFlags(&seq.ev, hipEventDisableTiming), "event")) return false;
        for (size_t i = 0; i < count; ++i)
            if (plans[i].kind == GatePlan::kDevice) pack_plan(plans[i], (char*)seq.host + offs[i]);
        if (!hip_ok(launch_upload(seq.dev, seq.host_dev, bytes, stream), "sequence upload")) return false;
    }
    for (size_t i = 0; i < count; ++i)
        if (!launch_plan(plans[i], plans[i].kind == GatePlan::kDevice ? (const char*)seq.dev + offs[i] : nullptr))
            return false;
    if (bytes) {
        if (!hip_ok(hipEventRecord(seq.ev, stream), "event record")) return false;
        seq.pending = true;
    }
    return true;
}

bool Engine::sync() { return hip_ok(hipStreamSynchronize(stream), "stream synchronize"); }

bool Engine::ensure_work(size_t doubles) {
    if (work_doubles >= doubles) return true;
    if (work) hipFree(work);
    work = nullptr;
    work_doubles = 0;
    if (hipMalloc(&work, doubles * sizeof(double)) != hipSuccess) {
        set_error(kErrOutOfMemory, "cannot allocate %zu scratch doubles", doubles);
        return false;
    }
    work_doubles = doubles;
    return true;
}

void Engine::trim_work(size_t keep_doubles) {
    if (work_doubles <= keep_doubles) return;
    sync();
    hipFree(work);
    work = nullptr;
    work_doubles = 0;
}

bool Engine::ensure_stage() {
    if (stage) return true;
    if (hipMalloc(&stage, 2 * kStage * sizeof(double)) != hipSuccess) {
        set_error(kErrOutOfMemory, "cannot allocate export staging");
        stage = nullptr;
        return false;
    }
    return true;
}

bool Engine::ensure_pinned(int arrays) { return ensure_pinned_doubles(2 * (size_t)arrays * std::min<uint64_t>(kStage, dim)); }

bool Engine::ensure_pinned_doubles(size_t need) {
    if (!pinned_ev[0])
        for (hipEvent_t& ev : pinned_ev)
            if (!hip_ok(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event")) return false;
    if (pinned_doubles >= need) return true;
    if (pinned) hipHostFree(pinned);
    pinned = nullptr;
    pinned_doubles = 0;
    if (hipHostMalloc((void**)&pinned, need * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        set_error(kErrOutOfMemory, "cannot allocate %zu pinned staging doubles", need);
        pinned = nullptr;
        return false;
    }
    pinned_doubles = need;
    return true;
}

namespace {
// memcpy split over host threads (page-aligned parts) for large copies
void parallel_copy(void* dst, const void* src, size_t bytes) {
    constexpr size_t kPart = 2u << 20;
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({16, hw, (bytes + kPart - 1) / kPart});
    if (nt <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = ((bytes / nt) + 4095) & ~size_t(4095);
    auto copy_part = [=](size_t t) {
        std::memcpy(static_cast<char*>(dst) + t * part, static_cast<const char*>(src) + t * part,
                    std::min(part, bytes - t * part));
    };
    std::vector<std::thread> pool;
    size_t t = 1;
    try {
        for (; t < nt && t * part < bytes; ++t) pool.emplace_back(copy_part, t);
    } catch (const std::exception&) {  // no thread to be had: copy the rest here
        for (; t < nt && t * part < bytes; ++t) copy_part(t);
    }
    copy_part(0);
    for (auto& th : pool) th.join();
}
}  // namespace

// Chunk c of kStage amplitudes: `produce` writes `arrays` device arrays at
// stage + a * kStage, a DMA moves them into pinned slot c % 2, and the host
// copies them out to dst[a] while the next chunk is produced and moved.
bool Engine::export_chunks(int arrays, double* const* dst, bool (*produce)(Engine*, uint64_t, uint64_t)) {
    if (!ensure_stage() || !ensure_pinned(arrays)) return false;
    const uint64_t per = std::min<uint64_t>(kStage, dim), nch = (dim + kStage - 1) / kStage;
    auto slot = [&](uint64_t c, int a) { return pinned + ((c & 1) * arrays + a) * per; };
    auto enqueue = [&](uint64_t c) {
        const uint64_t first = c * kStage, cnt = std::min<uint64_t>(kStage, dim - first);
        if (!produce(this, first, cnt)) return false;
        for (int a = 0; a < arrays; ++a)
            if (!hip_ok(hipMemcpyAsync(slot(c, a), stage + a * kStage, cnt * 8, hipMemcpyDeviceToHost, stream), "D2H"))
                return false;
        return hip_ok(hipEventRecord(pinned_ev[c & 1], stream), "event");
    };
    if (!enqueue(0)) return false;
    for (uint64_t c = 0; c < nch; ++c) {
        // slot (c + 1) % 2 was last read by chunk c - 1's host copy, which has returned
        if (c + 1 < nch && !enqueue(c + 1)) return false;
        if (!hip_ok(hipEventSynchronize(pinned_ev[c & 1]), "event")) return false;
        const uint64_t first = c * kStage, cnt = std::min<uint64_t>(kStage, dim - first);
        for (int a = 0; a < arrays; ++a) parallel_copy(dst[a] + first, slot(c, a), cnt * 8);
    }
    return sync();
}

bool Engine::download(const double* dsrc, double* hdst, uint64_t count) {
    if (count <= kStage) {
        return hip_ok(hipMemcpyAsync(hdst, dsrc, count * 8, hipMemcpyDeviceToHost, stream), "D2H") && sync();
    }
    if (!ensure_pinned_doubles(2 * kStage)) return false;
    const uint64_t per = pinned_doubles / 2, nch = (count + per - 1) / per;
    auto enqueue = [&](uint64_t c) {
        const uint64_t first = c * per, cnt = std::min<uint64_t>(per, count - first);
        return hip_ok(hipMemcpyAsync(pinned + (c & 1) * per, dsrc + first, cnt * 8, hipMemcpyDeviceToHost, stream),
                      "D2H") &&
               hip_ok(hipEventRecord(pinned_ev[c & 1], stream), "event");
    };
    if (!enqueue(0)) return false;
    for (uint64_t c = 0; c < nch; ++c) {
        if (c + 1 < nch && !enqueue(c + 1)) return false;
        if (!hip_ok(hipEventSynchronize(pinned_ev[c & 1]), "event")) return false;
        const uint64_t first = c * per, cnt = std::min<uint64_t>(per, count - first);
        parallel_copy(hdst + first, pinned + (c & 1) * per, cnt * 8);
    }
    return sync();
}

bool Engine::export_split(double* re, double* im) {
    double* const dst[2] = {re, im};
    return export_chunks(2, dst, [](Engine* e, uint64_t first, uint64_t cnt) {
        return hip_ok(launch_split(e->amp, first, cnt, e->prec, e->stage, e->stage + kStage, e->stream), "split");
    });
}

// The amplitudes as interleaved FP64 (re, im) pairs: chunk c is written into
// the stage by a kernel, moved to pinned slot c % 2, and copied out by host
// threads while chunk c + 1 moves.
bool Engine::export_interleaved(double* out) {
    if (!ensure_stage() || !ensure_pinned(2)) return false;
    const uint64_t per = std::min<uint64_t>(kStage, dim), nch = (dim + per - 1) / per;
    auto slot = [&](uint64_t c) { return pinned + (c & 1) * 2 * per; };
    auto enqueue = [&](uint64_t c) {
        const uint64_t first = c * per, cnt = std::min<uint64_t>(per, dim - first);
        return hip_ok(launch_interleave(amp, first, cnt, prec, stage, stream), "interleave") &&
               hip_ok(hipMemcpyAsync(slot(c), stage, cnt * 16, hipMemcpyDeviceToHost, stream), "D2H") &&
               hip_ok(hipEventRecord(pinned_ev[c & 1], stream), "event");
    };
    if (!enqueue(0)) return false;
    for (uint64_t c = 0; c < nch; ++c) {
        if (c + 1 < nch && !enqueue(c + 1)) return false;
        if (!hip_ok(hipEventSynchronize(pinned_ev[c & 1]), "event")) return false;
        const uint64_t first = c * per, cnt = std::min<uint64_t>(per, dim - first);
        parallel_copy(out + 2 * first, slot(c), cnt * 16);
    }
    return sync();
}

bool Engine::import_interleaved(const double* in) {
    if (!ensure_stage() || !ensure_pinned(2) || !sync()) return false;
    const uint64_t per = std::min<uint64_t>(kStage, dim), nch = (dim + per - 1) / per;
    for (uint64_t c = 0; c < nch; ++c) {
        const uint64_t first = c * per, cnt = std::min<uint64_t>(per, dim - first);
        double* ps = pinned + (c & 1) * 2 * per;
        if (c >= 2 && !hip_ok(hipEventSynchronize(pinned_ev[c & 1]), "event")) return false;
        parallel_copy(ps, in + 2 * first, cnt * 16);
        if (!hip_ok(hipMemcpyAsync(stage, ps, cnt * 16, hipMemcpyHostToDevice, stream), "H2D") ||
            !hip_ok(launch_deinterleave(amp, first, cnt, prec, stage, stream), "deinterleave") ||
            !hip_ok(hipEventRecord(pinned_ev[c & 1], stream), "event"))
            return false;
    }
    return sync();
}

bool Engine::export_probs(double* out) {
    double* const dst[1] = {out};
    return export_chunks(1, dst, [](Engine* e, uint64_t first, uint64_t cnt) {
        const void* src = static_cast<const char*>(e->amp) + first * e->elem;
        return hip_ok(launch_probs(src, cnt, e->prec, e->stage, e->stream), "probs");
    });
}

// Chunk c: the host copies re/im into pinned slot c % 2 (after that slot's
// previous DMA has finished), a DMA moves it to the device stage, and the join
// kernel writes the amplitudes, overlapping the next chunk's host copy.
bool Engine::import_split(const double* re, const double* im) {
    if (!ensure_stage() || !ensure_pinned(2) || !sync()) return false;
    const uint64_t per = std::min<uint64_t>(kStage, dim), nch = (dim + kStage - 1) / kStage;
    for (uint64_t c = 0; c < nch; ++c) {
        const uint64_t first = c * kStage, cnt = std::min<uint64_t>(kStage, dim - first);
        double* pr = pinned + (c & 1) * 2 * per;
        double* pi = pr + per;
        if (c >= 2 && !hip_ok(hipEventSynchronize(pinned_ev[c & 1]), "event")) return false;
        parallel_copy(pr, re + first, cnt * 8);
        parallel_copy(pi, im + first, cnt * 8);
        if (!hip_ok(hipMemcpyAsync(stage, pr, cnt * 8, hipMemcpyHostToDevice, stream), "H2D") ||
            !hip_ok(hipMemcpyAsync(stage + kStage, pi, cnt * 8, hipMemcpyHostToDevice, stream), "H2D") ||
            !hip_ok(launch_join(amp, first, cnt, prec, stage, stage + kStage, stream), "join") ||
            !hip_ok(hipEventRecord(pinned_ev[c & 1], stream), "event"))
            return false;
    }
    return sync();
}

bool Engine::fetch(int m, double* out) {
    if (!hip_ok(hipMemcpyAsync(host_out, red_out, m * sizeof(double), hipMemcpyDeviceToHost, stream), "D2H") || !sync())
        return false;
    for (int i = 0; i < m; ++i) out[i] = host_out[i];
    return true;
}

bool Engine::norm2(double* out) {
    return hip_ok(reduce_norm(amp, dim, prec, red_part, red_out, stream), "norm") && fetch(1, out);
}

bool Engine::bit_probs(int target, double* p01) {
    return hip_ok(reduce_bit_probs(amp, n, prec, target, red_part, red_out, stream), "bit probability") && fetch(2, p01);
}

bool Engine::pauli_terms(size_t nterms, const PauliTerm* terms, double* vals) {
    // group strings by flip mask, up to 8 per pass
    std::vector<int> done(nterms, 0);
    for (size_t i = 0; i < nterms; ++i) {
        if (done[i]) continue;
        std::vector<size_t> grp;
        for (size_t j = i; j < nterms && grp.size() < 8; ++j)
            if (!done[j] && terms[j].xmask == terms[i].xmask) {
                grp.push_back(j);
                done[j] = 1;
            }
        uint64_t zy[8];
        int ny[8];
        for (size_t s = 0; s < grp.size(); ++s) {
            zy[s] = terms[grp[s]].zymask;
            ny[s] = terms[grp[s]].ny;
        }
        double tmp[16];
        if (!hip_ok(reduce_pauli_group(amp, n, prec, terms[i].xmask, (int)grp.size(), zy, ny, red_part, red_out, stream),
                    "pauli expectation") ||
            !fetch(2 * (int)grp.size(), tmp))
            return false;
        for (size_t s = 0; s < grp.size(); ++s) {
            vals[2 * grp[s]] = tmp[2 * s];
            vals[2 * grp[s] + 1] = tmp[2 * s + 1];
        }
    }
    return true;
}

bool parse_pauli(const char* ops, const size_t* qubits, size_t len, int n, PauliTerm& t) {
    t.xmask = 0;
    t.zymask = 0;
    t.ny = 0;
    for (size_t k = 0; k < len; ++k) {
        const size_t q = qubits[k];
        const char c = ops[k];
        if (c != 'X' && c != 'Y' && c != 'Z') continue;  // other characters act as identity
        if (q >= (size_t)n) {
            set_error(kErrInvalidArgument, "Pauli qubit %zu out of range for %d qubits", q, n);
            return false;
        }
        const uint64_t b = 1ULL << q;
        if (c == 'X') t.xmask ^= b;
        else if (c == 'Y') {
            t.xmask ^= b;
            t.zymask ^= b;
            t.ny += 1;
        } else t.zymask ^= b;
    }
    return true;
}


}  // namespace iqg

// ------------------------------------------------------------------ C ABI

namespace {

Engine* sv_of(void* p) { return static_cast<Engine*>(p); }

struct DMHandle {
    Engine eng;  // 2n-qubit vectorised rho, complex128
    int nq = 0;
};

DMHandle* dm_of(void* p) { return static_cast<DMHandle*>(p); }

void to_u32(const size_t* q, size_t k, uint32_t* out) {
    for (size_t i = 0; i < k; ++i) out[i] = (uint32_t)std::min<size_t>(q[i], 0xffffffffu);
}

// S = sum_k K (x) conj(K), both d x d; S is (d*d) x (d*d) row-major (MSB-first: row part high).
void superop(size_t d, size_t nk, const double* kre, const double* kim, std::vector<double>& sre, std::vector<double>& sim) {
    const size_t D = d * d;
    sre.assign(D * D, 0.0);
    sim.assign(D * D, 0.0);
    for (size_t k = 0; k < nk; ++k) {
        const double* ar = kre + k * d * d;
        const double* ai = kim + k * d * d;
        for (size_t r = 0; r < d; ++r)
            for (size_t c = 0; c < d; ++c)
                for (size_t r2 = 0; r2 < d; ++r2)
                    for (size_t c2 = 0; c2 < d; ++c2) {
                        // K[r, r2] * conj(K[c, c2])
                        const double xr = ar[r * d + r2], xi = ai[r * d + r2];
                        const double yr = ar[c * d + c2], yi = -ai[c * d + c2];
                        const size_t row = r * d + c, col = r2 * d + c2;
                        sre[row * D + col] += xr * yr - xi * yi;
                        sim[row * D + col] += xr * yi + xi * yr;
                    }
    }
}

bool dm_apply_superop(DMHandle* h, const uint32_t* q, int k, size_t nk, const double* kre, const double* kim) {
    std::vector<double> sre, sim;
    superop((size_t)1 << k, nk, kre, kim, sre, sim);
    uint32_t bits[4];
    for (int i = 0; i < k; ++i) {
        bits[i] = q[i] + (uint32_t)h->nq;  // row qubits
        bits[k + i] = q[i];                 // column qubits
    }
    return h->eng.apply(2 * k, bits, sre.data(), sim.data(), gate_path_override());
}

}  // namespace

extern "C" {

int iq_last_error(void) { return t_err; }
const char* iq_last_error_message(void) { return t_msg; }
void iq_clear_error(void) {
    t_err = kOk;
    t_msg[0] = '\0';
}
const char* iq_backend_name(void) { return "hip-gfx950"; }
void iq_set_gate_path(int path) { g_gate_path = (path >= 0 && path <= 3) ? path : 0; }
void iq_set_mfma_options(int batches_per_iteration, int matrix_in_lds, int nontemporal, int grid_waves,
                        int batch_rotation) {
    set_mfma_options(batches_per_iteration, matrix_in_lds, nontemporal, grid_waves, batch_rotation);
}

// ------------------------------------------------------------- statevector

void* iq_statevector_create_precision(size_t num_qubits, int precision) {
    if (num_qubits > (size_t)kMaxN || (precision != kC128 && precision != kC64)) {
        set_error(kErrInvalidArgument, "unsupported statevector shape (n=%zu, precision=%d)", num_qubits, precision);
        return nullptr;
    }
    Engine* e = new (std::nothrow) Engine();
    if (!e) return nullptr;
    if (!e->init((int)num_qubits, precision)) {
        delete e;
        return nullptr;
    }
    return e;
}

void* iq_statevector_create(size_t num_qubits) { return iq_statevector_create_precision(num_qubits, kC128); }

int iq_statevector_get_precision(void* sv) { return sv ? sv_of(sv)->prec : -1; }

void iq_statevector_free(void* sv) {
    if (!sv) return;
    DeviceGuard g(sv_of(sv)->dev);
    delete sv_of(sv);
}

void iq_statevector_reset(void* sv) {
    if (!sv) return;
    DeviceGuard g(sv_of(sv)->dev);
    sv_of(sv)->reset_basis(0);
}

void iq_statevector_apply_1q(void* sv, size_t target, const double* u_real, const double* u_imag) {
    if (!sv || !u_real || !u_imag) return;
    DeviceGuard g(sv_of(sv)->dev);
    uint32_t q = (uint32_t)std::min<size_t>(target, 0xffffffffu);
    sv_of(sv)->apply(1, &q, u_real, u_imag, gate_path_override());
}

void iq_statevector_apply_2q(void* sv, size_t q0, size_t q1, const double* u_real, const double* u_imag) {
    if (!sv || !u_real || !u_imag) return;
    DeviceGuard g(sv_of(sv)->dev);
    const size_t qs[2] = {q0, q1};
    uint32_t q[2];
    to_u32(qs, 2, q);
    sv_of(sv)->apply(2, q, u_real, u_imag, gate_path_override());
}

void iq_statevector_apply_fused(void* sv, const size_t* qubits, size_t k, const double* u_real, const double* u_imag) {
    if (!sv || !qubits || !u_real || !u_imag) return;
    DeviceGuard g(sv_of(sv)->dev);
    if (k < 1 || k > (size_t)kMaxK) {
        set_error(kErrInvalidArgument, "fused gate on %zu qubits is not supported (1..5)", k);
        return;
    }
    uint32_t q[kMaxK];
    to_u32(qubits, k, q);
    sv_of(sv)->apply((int)k, q, u_real, u_imag, gate_path_override());
}

// A sequence of fused gates in one call (the output of a fusion pass): gate i
// acts on ks[i] qubits (1..5) taken in order from `qubits`, its 4^ks[i]
// matrix entries in order from u_real / u_imag, MSB-first as apply_fused.
// All gates are checked before any runs; their matrices go up in one copy.
void iq_statevector_apply_fused_sequence(void* sv, size_t count, const size_t* ks, const size_t* qubits,
                                         const double* u_real, const double* u_imag) {
    if (!sv || !ks || !qubits || !u_real || !u_imag || count == 0) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    std::vector<int> k32(count);
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) {
        if (ks[i] < 1 || ks[i] > (size_t)kMaxK) {
            set_error(kErrInvalidArgument, "fused gate %zu on %zu qubits is not supported (1..5)", i, ks[i]);
            return;
        }
        k32[i] = (int)ks[i];
        total += ks[i];
    }
    std::vector<uint32_t> q(total);
    to_u32(qubits, total, q.data());
    e->apply_sequence(count, k32.data(), q.data(), u_real, u_imag, gate_path_override());
}

void iq_statevector_apply_diagonal(void* sv, const size_t* qubits, size_t k, const double* d_real, const double* d_imag) {
    if (!sv || !qubits || !d_real || !d_imag) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    if (k < 1 || k > (size_t)kMaxK || k > (size_t)e->n) {
        set_error(kErrInvalidArgument, "diagonal gate on %zu qubits is not supported", k);
        return;
    }
    uint32_t q[kMaxK];
    to_u32(qubits, k, q);
    for (size_t i = 0; i < k; ++i)
        if ((int)q[i] >= e->n) {
            set_error(kErrInvalidArgument, "qubit index out of range");
            return;
        }
    hip_ok(launch_diag(e->amp, e->n, e->prec, e->stream, (int)k, q, d_real, d_imag), "diagonal gate launch");
}

double iq_statevector_measure_prob(void* sv, size_t target, int outcome) {
    if (!sv) return 0.0;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    if (target >= (size_t)e->n) {
        set_error(kErrInvalidArgument, "target qubit %zu out of range", target);
        return 0.0;
    }
    double p[2];
    if (!e->bit_probs((int)target, p)) return 0.0;
    return (outcome == 0 || outcome == 1) ? p[outcome] : 0.0;
}

// reference core/src/statevector.cpp:171-195: outcome = rand < P0 ? 0 : 1,
// renormalise by 1/sqrt(P(outcome)); P(outcome) < 1e-15 is an error.
int iq_statevector_measure(void* sv, size_t target, double rand_val) {
    if (!sv) return 0;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    if (target >= (size_t)e->n) {
        set_error(kErrInvalidArgument, "target qubit %zu out of range", target);
        return 0;
    }
    double p[2];
    if (!e->bit_probs((int)target, p)) return 0;
    const int outcome = (rand_val < p[0]) ? 0 : 1;
    if (p[outcome] < 1e-15) {
        set_error(kErrZeroProbability, "attempted to collapse to a zero-probability state");
        return outcome;
    }
    hip_ok(launch_collapse(e->amp, e->n, e->prec, (int)target, outcome, 1.0 / std::sqrt(p[outcome]), e->stream), "collapse");
    return outcome;
}

void iq_statevector_reset_qubit(void* sv, size_t target, double rand_val) {
    if (!sv) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    if (target >= (size_t)e->n) {
        set_error(kErrInvalidArgument, "target qubit %zu out of range", target);
        return;
    }
    const int before = t_err;
    iq_clear_error();
    const int outcome = iq_statevector_measure(sv, target, rand_val);
    if (t_err != kOk) return;
    t_err = before;
    if (outcome == 1) {
        const double xr[4] = {0, 1, 1, 0}, xi[4] = {0, 0, 0, 0};
        uint32_t q = (uint32_t)target;
        e->apply(1, &q, xr, xi, gate_path_override());
    }
}

void iq_statevector_get_amplitudes(void* sv, double* real_out, double* imag_out) {
    if (!sv || !real_out || !imag_out) return;
    DeviceGuard g(sv_of(sv)->dev);
    sv_of(sv)->export_split(real_out, imag_out);
}

void iq_statevector_set_amplitudes(void* sv, const double* real_in, const double* imag_in) {
    if (!sv || !real_in || !imag_in) return;
    DeviceGuard g(sv_of(sv)->dev);
    sv_of(sv)->import_split(real_in, imag_in);
}

// Interleaved (re, im) FP64 copies of the amplitudes, 2 * 2^n doubles (no
// reference counterpart: c_api.h:20 exports split arrays).
void iq_statevector_get_amplitudes_interleaved(void* sv, double* out) {
    if (!sv || !out) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    e->export_interleaved(out);
}

void iq_statevector_set_amplitudes_interleaved(void* sv, const double* in) {
    if (!sv || !in) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    e->import_interleaved(in);
}

void iq_statevector_get_probabilities(void* sv, double* probs_out) {
    if (!sv || !probs_out) return;
    DeviceGuard g(sv_of(sv)->dev);
    sv_of(sv)->export_probs(probs_out);
}

double iq_statevector_norm_squared(void* sv) {
    if (!sv) return 0.0;
    DeviceGuard g(sv_of(sv)->dev);
    double v = 0.0;
    sv_of(sv)->norm2(&v);
    return v;
}

void iq_statevector_normalize(void* sv) {
    if (!sv) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    double v = 0.0;
    if (!e->norm2(&v) || !(v > 0.0)) return;
    hip_ok(launch_scale(e->amp, e->dim, e->prec, 1.0 / std::sqrt(v), e->stream), "scale");
}

// Sampling buffers in the work area: the CDF (when it is written), the scan
// scratch, the uniforms and the results.
namespace {
struct SampleBufs {
    bool direct;  // per-chunk searches, no CDF written
    double* cdf;
    double* scratch;
    double* du;
    uint64_t* dout;
};

static bool sample_bufs(Engine* e, size_t shots, SampleBufs& b) {
    b.direct = sample_without_cdf(e->dim, shots);
    const size_t ncdf = b.direct ? 0 : e->dim, nscr = scan_scratch_doubles(e->dim), ns = std::max<size_t>(shots, 1);
    if (!e->ensure_work(ncdf + nscr + 2 * ns)) return false;
    b.cdf = e->work;
    b.scratch = e->work + ncdf;
    b.du = b.scratch + nscr;
    b.dout = reinterpret_cast<uint64_t*>(b.du + ns);
    return true;
}

// the passes over the state that do not need the uniforms
static bool sample_prepare(Engine* e, const SampleBufs& b) {
    return b.direct ? hip_ok(launch_cdf_chunks(e->amp, e->dim, e->prec, b.scratch, e->stream), "cdf")
                    : hip_ok(launch_cdf(e->amp, e->dim, e->prec, b.cdf, b.scratch, e->stream), "cdf");
}

// uniforms in b.du -> indices in b.dout
static bool sample_search(Engine* e, const SampleBufs& b, size_t shots) {
    return b.direct
               ? hip_ok(launch_sample_chunks(e->amp, e->dim, e->prec, b.scratch, b.du, shots, b.dout, e->stream), "search")
               : hip_ok(launch_search(b.cdf, e->dim, b.du, shots, b.dout, e->stream), "search");
}
}  // namespace

// reference core/src/statevector.cpp:235-256: probabilities, inclusive CDF,
// uniform_real_distribution<double>(0,1) over mt19937_64(seed), lower_bound,
// clamp to dim-1.  The uniforms are drawn here on the host with the same
// standard-library engine and distribution; the CDF and the searches run on the GPU.
void iq_statevector_sample(void* sv, size_t num_shots, uint64_t seed, uint64_t* results_out) {
    if (!sv || !results_out) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    SampleBufs b;
    if (!sample_bufs(e, num_shots, b)) return;
    std::vector<double> u(num_shots);
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    for (size_t s = 0; s < num_shots; ++s) u[s] = uni(gen);
    if (!sample_prepare(e, b)) return;
    if (num_shots == 0) {
        e->sync();
        return;
    }
    if (!hip_ok(hipMemcpyAsync(b.du, u.data(), num_shots * 8, hipMemcpyHostToDevice, e->stream), "H2D") ||
        !sample_search(e, b, num_shots) ||
        !hip_ok(hipMemcpyAsync(results_out, b.dout, num_shots * 8, hipMemcpyDeviceToHost, e->stream), "D2H"))
        return;
    e->sync();
}

// Several sampling tasks on one state in one call (the tasks of a distributed
// job, invariantq/distributed/planner.py:110-163, each of which the reference
// serves by re-simulating the circuit, invariantq/distributed/worker.py:52-95):
// task t draws shots[t] indices with std::mt19937_64(seeds[t]) exactly as
// iq_statevector_sample(sv, shots[t], seeds[t], ...) does; results are
// concatenated in task order.  The passes over the state run once and one
// search launch serves every shot; the tasks' uniforms are drawn on host
// threads meanwhile.
void iq_statevector_sample_tasks(void* sv, size_t ntasks, const uint64_t* seeds, const uint64_t* shots,
                                 uint64_t* results_out) {
    if (!sv || !results_out || !seeds || !shots || ntasks == 0) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    std::vector<size_t> off(ntasks + 1, 0);
    for (size_t t = 0; t < ntasks; ++t) {
        if (shots[t] > (SIZE_MAX / 16) - off[t]) {
            set_error(kErrInvalidArgument, "total shot count overflows");
            return;
        }
        off[t + 1] = off[t] + shots[t];
    }
    const size_t total = off[ntasks];
    SampleBufs b;
    if (!sample_bufs(e, total, b) || !sample_prepare(e, b)) return;
    std::vector<double> u(total);
    auto draw = [&](size_t t0, size_t t1) {
        for (size_t t = t0; t < t1; ++t) {
            std::mt19937_64 gen(seeds[t]);
            std::uniform_real_distribution<double> uni(0.0, 1.0);
            for (size_t s = off[t]; s < off[t + 1]; ++s) u[s] = uni(gen);
        }
    };
    const size_t nthreads = std::min<size_t>({ntasks, 16, std::max<size_t>(1, total / 65536)});
    if (nthreads <= 1) {
        draw(0, ntasks);
    } else {
        std::vector<std::thread> pool;
        for (size_t i = 0; i < nthreads; ++i) pool.emplace_back(draw, ntasks * i / nthreads, ntasks * (i + 1) / nthreads);
        for (auto& th : pool) th.join();
    }
    if (total == 0) {
        e->sync();
        return;
    }
    if (!hip_ok(hipMemcpyAsync(b.du, u.data(), total * 8, hipMemcpyHostToDevice, e->stream), "H2D") ||
        !sample_search(e, b, total) ||
        !hip_ok(hipMemcpyAsync(results_out, b.dout, total * 8, hipMemcpyDeviceToHost, e->stream), "D2H"))
        return;
    e->sync();
}

// Device-side uniforms variant: the caller passes the uniforms (e.g. from a
// seed tree); lower_bound semantics as iq_statevector_sample.
void iq_statevector_sample_uniforms(void* sv, size_t num_shots, const double* uniforms, uint64_t* results_out) {
    if (!sv || !results_out || (!uniforms && num_shots)) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    SampleBufs b;
    if (!sample_bufs(e, num_shots, b) || !sample_prepare(e, b)) return;
    if (num_shots) {
        if (!hip_ok(hipMemcpyAsync(b.du, uniforms, num_shots * 8, hipMemcpyHostToDevice, e->stream), "H2D") ||
            !sample_search(e, b, num_shots) ||
            !hip_ok(hipMemcpyAsync(results_out, b.dout, num_shots * 8, hipMemcpyDeviceToHost, e->stream), "D2H"))
            return;
    }
    e->sync();
}

// ------------------------------------------------ sampling without allocation
// Scan scratch of the per-chunk sampling path (chunk offsets and running
// maxima, group offsets and maxima) for states of up to 2^31 amplitudes:
// 2 * 2^19 + 2 * 2^7 doubles, 8 MiB per GPU, held until the process ends.
// Every entry is written by the scan before it is read, so nothing needs
// clearing between calls.
constexpr uint64_t kNoAllocMaxAmps = 1ULL << 31;
constexpr size_t kNoAllocScratchDoubles = 2 * (kNoAllocMaxAmps / 4096) + 2 * (kNoAllocMaxAmps / 4096 / 4096);
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

// (static: these helpers sit inside the extern "C" block)
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's `scratch` (the host-side name is not one),
// looked up once per device with that device current.
static double* noalloc_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) {
        set_error(kErrUnsupported, "device index %d beyond %d", dev, kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address")) return nullptr;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

static bool device_accessible(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // an unregistered host pointer leaves an error behind
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
           (a.type == hipMemoryTypeHost && a.devicePointer != nullptr);
}

// As iq_statevector_sample_uniforms (the same indices for the same uniforms),
// with uniforms and results in device memory, nothing allocated and nothing
// waited for: the scan and the searches are enqueued on the handle's stream,
// whose completion makes the results valid.  Always the per-chunk searches
// (no CDF written).  States above 2^31 amplitudes are rejected.  Every call
// on one GPU shares that GPU's scratch: calls that overlap on one GPU (other
// handles' streams, other host threads) race.
void iq_statevector_sample_uniforms_noalloc(void* sv, size_t num_shots, const double* uniforms,
                                            uint64_t* results_out) {
    if (!sv) return;
    if (num_shots == 0) return;
    if (!uniforms || !results_out) {
        set_error(kErrInvalidArgument, "null uniforms or results");
        return;
    }
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    if (e->dim > kNoAllocMaxAmps) {
        set_error(kErrUnsupported, "sampling without allocation takes states of at most 2^31 amplitudes");
        return;
    }
    if (!device_accessible(uniforms) || !device_accessible(results_out)) {
        set_error(kErrInvalidArgument, "uniforms and results must be device-accessible memory");
        return;
    }
    if (scan_scratch_doubles(e->dim) > kNoAllocScratchDoubles) {
        set_error(kErrUnsupported, "scan scratch beyond the static buffer");
        return;
    }
    double* sc = noalloc_scratch(e->dev);
    if (!sc) return;
    if (!hip_ok(launch_cdf_chunks(e->amp, e->dim, e->prec, sc, e->stream), "cdf")) return;
    hip_ok(launch_sample_chunks(e->amp, e->dim, e->prec, sc, uniforms, num_shots, results_out, e->stream), "search");
}

// Inclusive FP64 CDF of the probabilities (the running maximum of the chunked
// prefix sums, so monotone), exported for boundary-margin checks; sampling
// returns the first index whose value here reaches the uniform.
void iq_statevector_get_cdf(void* sv, double* cdf_out) {
    if (!sv || !cdf_out) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    if (!e->ensure_work(e->dim + scan_scratch_doubles(e->dim) + 2)) return;
    // through the pinned slots (a pageable copy ran at about 10 GB/s)
    if (!hip_ok(launch_cdf(e->amp, e->dim, e->prec, e->work, e->work + e->dim, e->stream), "cdf")) return;
    e->download(e->work, cdf_out, e->dim);
}

// Reduced density matrices of up to 15 qubits: 2^30 complex doubles (16 GiB)
// on the device and again on the host.
constexpr size_t kMaxReducedQubits = 15;

// Reduced density matrix of `qubits` (bit b of a row index = qubits[b]) in
// the work buffer (kd x kd double2 at its start), with at least
// `extra_doubles` of the buffer free after it: the partial trace of reference
// core/src/statevector.cpp:293-330, on the matrix cores for 4 or more kept and
// 2 or more traced qubits.
static bool reduced_matrix_device(Engine* e, const size_t* qubits, size_t k, size_t extra_doubles) {
    if (k < 1 || k > kMaxReducedQubits || k > (size_t)e->n) {
        set_error(k > kMaxReducedQubits ? kErrUnsupported : kErrInvalidArgument,
                  "reduced matrices of 1..15 qubits (at most the state's) are supported");
        return false;
    }
    uint32_t q[16];
    to_u32(qubits, k, q);
    for (size_t i = 0; i < k; ++i) {
        if ((int)q[i] >= e->n) {
            set_error(kErrInvalidArgument, "subsystem qubit out of range");
            return false;
        }
        for (size_t j = 0; j < i; ++j)
            if (q[i] == q[j]) {
                set_error(kErrInvalidArgument, "repeated subsystem qubit");
                return false;
            }
    }
    const size_t kd = (size_t)1 << k;
    // reduced matrix, then the partial sums of the environment chunks and
    // (matrix-core path, unless the kept qubits are 0..k-1 of a complex128
    // state) the state reordered to (environment, kept)
    bool in_order = e->prec == kC128;
    for (size_t i = 0; i < k; ++i) in_order &= q[i] == (uint32_t)i;
    const bool gather = k >= 4 && (size_t)e->n - k >= 2 && !in_order;
    const size_t parts = k <= 3   ? 2048 * (kd * (kd + 1) / 2)
                         : k <= 5  ? 2048 * kd * kd
                         : k <= 10 ? std::min<size_t>(1u << 22, 512 * kd * kd)
                                   : 0;
    const size_t after = 2 * ((gather ? e->dim : 0) + parts);
    if (!e->ensure_work(2 * kd * kd + std::max<size_t>(after, extra_doubles))) return false;
    double2* out = reinterpret_cast<double2*>(e->work);
    return hip_ok(launch_partial_trace(e->amp, e->n, e->prec, q, (int)k, out, out + kd * kd,
                                       (e->work_doubles - 2 * kd * kd) / 2, e->stream),
                  "partial trace");
}

// The same in a host buffer of 2 * 4^k doubles (re, im interleaved; not
// zero-filled first).
static bool reduced_matrix(Engine* e, const size_t* qubits, size_t k, std::unique_ptr<double[]>& red) {
    if (!reduced_matrix_device(e, qubits, k, 0)) return false;
    const size_t kd = (size_t)1 << k;
    red.reset(new (std::nothrow) double[2 * kd * kd]);
    if (!red) {
        set_error(kErrOutOfMemory, "cannot allocate %zu host doubles for the reduced matrix", 2 * kd * kd);
        return false;
    }
    return e->download(e->work, red.get(), 2 * kd * kd);
}

// jacobi_rows.hpp's backend on the engine's device: the matrix at the start
// of the work buffer, the row list, the records and the shared candidate list
// after it, with the same layout in the pinned staging buffer.  Each rotation
// is one round trip: the row list up, the rotation and the scans, then one
// copy down of the block that holds the new columns and the row records (and,
// when rows have more than one candidate, a second of the shared list).
struct DeviceRows {
    static constexpr uint32_t kCap = 1u << 16;  // shared candidate list
    Engine* e = nullptr;
    double2* A = nullptr;
    uint32_t d = 0;
    unsigned char *dev = nullptr, *host = nullptr;  // same layout
    // fixed parts: the shared list (t, j, value), the row list; then a block
    // laid out per call (count of the shared list, row records, columns)
    size_t off_ev = 0, off_et = 0, off_ej = 0, off_rows = 0, off_blk = 0;

    static size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }
    static size_t aux_bytes(size_t d) {
        return 16 * (size_t)kCap + 8 * (size_t)kCap + align16(4 * d) + 16 + align16(8 * d) + 16 * d + 32 * d + 64;
    }
    static size_t aux_doubles(size_t d) { return (aux_bytes(d) + 7) / 8; }

    bool init(Engine* eng, uint32_t dim) {
        e = eng;
        d = dim;
        A = reinterpret_cast<double2*>(e->work);
        const size_t aux = aux_doubles(d);
        if (e->work_doubles < 2 * (size_t)d * d + aux) {
            set_error(kErrOutOfMemory, "Jacobi records do not fit");
            return false;
        }
        if (!e->ensure_pinned_doubles(aux)) return false;
        dev = reinterpret_cast<unsigned char*>(e->work + 2 * (size_t)d * d);
        host = reinterpret_cast<unsigned char*>(e->pinned);
        off_ev = 0;
        off_et = off_ev + 16 * (size_t)kCap;
        off_ej = off_et + 4 * (size_t)kCap;
        off_rows = off_ej + 4 * (size_t)kCap;
        off_blk = off_rows + align16(4 * (size_t)d);
        return true;
    }
    static uint32_t* u32(unsigned char* base, size_t off) { return reinterpret_cast<uint32_t*>(base + off); }
    static double2* d2(unsigned char* base, size_t off) { return reinterpret_cast<double2*>(base + off); }
    static const std::complex<double>* cx(unsigned char* base, size_t off) {
        return reinterpret_cast<const std::complex<double>*>(base + off);
    }

    // the block for nrows records and q column entries: [count of the shared
    // list | counts | first columns | first values | column p | column q]
    struct Block {
        size_t en, cnt, fj, fv, colp, colq, end;
    };
    static Block block(size_t nrows, size_t q) {
        Block b;
        b.en = 0;
        b.cnt = 16;
        b.fj = b.cnt + 4 * nrows;
        b.fv = align16(b.fj + 4 * nrows);
        b.colp = b.fv + 16 * nrows;
        b.colq = b.colp + 16 * q;
        b.end = b.colq + 16 * q;
        return b;
    }

    // through the block area (a strided 8-byte copy of the diagonal took 57 ms
    // at d = 8192)
    bool diagonal(double* re) {
        if (!hip_ok(launch_diag_real(A, d, reinterpret_cast<double*>(dev + off_blk), e->stream), "diagonal") ||
            !hip_ok(hipMemcpyAsync(host + off_blk, dev + off_blk, (size_t)d * 8, hipMemcpyDeviceToHost, e->stream),
                    "D2H") ||
            !e->sync())
            return false;
        std::memcpy(re, host + off_blk, (size_t)d * 8);
        return true;
    }

    bool upload_rows(const uint32_t* rows, size_t nrows) {
        std::memcpy(host + off_rows, rows, nrows * 4);
        return hip_ok(hipMemcpyAsync(dev + off_rows, host + off_rows, nrows * 4, hipMemcpyHostToDevice, e->stream),
                      "H2D");
    }

    // scans of the listed rows into the block, then the block down and the wait
    bool scans(size_t nrows, const Block& b) {
        unsigned char* bd = dev + off_blk;
        return hip_ok(launch_jacobi_scan(A, d, u32(dev, off_rows), (uint32_t)nrows, u32(bd, b.cnt),
                                         u32(bd, b.fj), d2(bd, b.fv), u32(bd, b.en), kCap,
                                         u32(dev, off_et), u32(dev, off_ej), d2(dev, off_ev),
                                         e->stream),
                      "Jacobi scan") &&
               hip_ok(hipMemcpyAsync(host + off_blk, bd, b.end, hipMemcpyDeviceToHost, e->stream), "D2H") && e->sync();
    }

    // the shared list (when used) and the records' host view
    bool finish(const Block& b, RowScan& out) {
        unsigned char* bh = host + off_blk;
        const uint32_t en = *u32(bh, b.en), take = std::min(en, kCap);
        if (take &&
            !(hip_ok(hipMemcpyAsync(host + off_ev, dev + off_ev, take * 16, hipMemcpyDeviceToHost, e->stream), "D2H") &&
              hip_ok(hipMemcpyAsync(host + off_et, dev + off_et, take * 4, hipMemcpyDeviceToHost, e->stream), "D2H") &&
              hip_ok(hipMemcpyAsync(host + off_ej, dev + off_ej, take * 4, hipMemcpyDeviceToHost, e->stream), "D2H") &&
              e->sync()))
            return false;
        out.count = u32(bh, b.cnt);
        out.first_j = u32(bh, b.fj);
        out.first_v = cx(bh, b.fv);
        out.extra_n = en;
        out.extra_cap = kCap;
        out.extra_t = u32(host, off_et);
        out.extra_j = u32(host, off_ej);
        out.extra_v = cx(host, off_ev);
        return true;
    }

    bool scan(const uint32_t* rows, size_t nrows, RowScan& out) {
        const Block b = block(nrows, 0);
        return upload_rows(rows, nrows) &&
               hip_ok(hipMemsetAsync(dev + off_blk + b.en, 0, 4, e->stream), "memset") && scans(nrows, b) &&
               finish(b, out);
    }

    bool rotate_scan(size_t p, size_t q, double c, std::complex<double> sn, double npp, double nqq,
                     const std::complex<double>*& colp, const std::complex<double>*& colq, const uint32_t* rows,
                     size_t nrows, RowScan& out) {
        const Block b = block(nrows, q);
        unsigned char* bd = dev + off_blk;
        // the new columns of rows i < q go into the block
        if (!upload_rows(rows, nrows) ||
            !hip_ok(launch_jacobi_rotate(A, d, (uint32_t)p, (uint32_t)q, c, sn.real(), sn.imag(), npp, nqq,
                                         d2(bd, b.colp), d2(bd, b.colq), u32(bd, b.en),
                                         (uint32_t)q, e->stream),
                    "Jacobi rotation") ||
            !scans(nrows, b))
            return false;
        colp = cx(host + off_blk, b.colp);
        colq = cx(host + off_blk, b.colq);
        return finish(b, out);
    }

    bool fetch_row(size_t i, std::complex<double>* out) {
        const size_t len = (size_t)d - i - 1;
        return hip_ok(hipMemcpyAsync(out, A + i * d + i + 1, len * 16, hipMemcpyDeviceToHost, e->stream), "D2H") &&
               e->sync();
    }
};

// Kept qubits from which the entropy's Jacobi runs on the device (the matrix
// stays there).  Measured per call: 10 kept of 20 qubits 3.0-3.8 ms against
// 7.3-37 ms for the host's row-maximum Jacobi on a downloaded copy, 13 of 26
// 55 against 301 ms, 15 of 30 2.6 against 4.6 s; at 8 kept the host is ahead
// (2.8 against 3.3 ms: a rotation's device round trip is about 100 us).
constexpr int kDeviceJacobiDefault = 10;
static std::atomic<int> g_device_jacobi_min{kDeviceJacobiDefault};

// the reduced-matrix scratch is given back once it passes 2 GiB (13 or more
// kept qubits, or large gathered states), so one entropy call does not hold
// tens of GiB of device memory for the rest of the handle's life
constexpr size_t kKeepWorkDoubles = (size_t)1 << 28;

// reference core/src/statevector.cpp:395-410: entropy from the capped Jacobi
// eigenvalues of the reduced matrix.  0 for an empty or full subsystem, as
// the reference; NaN (with the error set) when the entropy cannot be computed
// (invalid qubits, more than 15 kept, no memory), never a plausible value.
double iq_statevector_entropy(void* sv, const size_t* qubits, size_t num_subsys_qubits) {
    if (!sv || !qubits) return 0.0;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    const size_t k = num_subsys_qubits;
    if (k == 0 || k == (size_t)e->n) return 0.0;
    if ((int)k >= g_device_jacobi_min.load() && k <= kMaxReducedQubits) {
        // rho_A stays on the device; the host keeps the row records
        const uint32_t kd = 1u << k;
        DeviceRows dev;
        std::vector<double> ev;
        const bool ok = reduced_matrix_device(e, qubits, k, DeviceRows::aux_doubles(kd)) && dev.init(e, kd) &&
                        jacobi_eigenvalues_rows(dev, kd, 100, 1e-12, ev);
        e->trim_work(kKeepWorkDoubles);
        if (!ok) return std::numeric_limits<double>::quiet_NaN();
        return entropy_bits(ev);
    }
    std::unique_ptr<double[]> red;
    const bool ok = reduced_matrix(e, qubits, k, red);
    e->trim_work(kKeepWorkDoubles);
    if (!ok) return std::numeric_limits<double>::quiet_NaN();
    auto* A = reinterpret_cast<std::complex<double>*>(red.get());
    return entropy_bits(jacobi_eigenvalues_inplace(A, (size_t)1 << k, 100, 1e-12));
}

// Kept qubits from which the entropy's Jacobi runs on the device (1..16; 16
// keeps every size on the host; 0 restores the default).  Both give the same
// result bit for bit.
void iq_set_entropy_options(int device_jacobi_min_kept) {
    g_device_jacobi_min.store(device_jacobi_min_kept <= 0 ? kDeviceJacobiDefault
                                                          : std::min(device_jacobi_min_kept, 16));
}

// Reduced density matrix export (no C export in the reference: partial_trace
// is a class method, core/src/statevector.cpp:293-330), 2^k x 2^k row-major.
void iq_statevector_reduced_density_matrix(void* sv, const size_t* qubits, size_t k, double* re_out,
                                           double* im_out) {
    if (!sv || !qubits || !re_out || !im_out) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    std::unique_ptr<double[]> red;
    const bool ok = reduced_matrix(e, qubits, k, red);
    e->trim_work(kKeepWorkDoubles);
    if (!ok) return;
    const size_t total = (size_t)1 << (2 * k);
    for (size_t i = 0; i < total; ++i) {
        re_out[i] = red[2 * i];
        im_out[i] = red[2 * i + 1];
    }
}

void iq_statevector_expval_pauli(void* sv, const char* pauli_str, const size_t* qubits, size_t len, double* val_real,
                                 double* val_imag) {
    if (!sv || !pauli_str || !qubits || !val_real || !val_imag) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    PauliTerm t;
    if (!parse_pauli(pauli_str, qubits, len, e->n, t)) return;
    double v[2] = {0.0, 0.0};
    if (!e->pauli_terms(1, &t, v)) return;
    *val_real = v[0];
    *val_imag = v[1];
}

// Sum of Pauli strings: term t has lens[t] characters of `ops` and qubits,
// consumed consecutively.  Writes sum_t coeff[t] <P_t> and, if per_term is
// non-null, the 2*nterms values Re/Im <P_t>.  Without per_term, sums over
// many flip masks on 20 to 34 qubits run over tile passes (a few reads of the
// state instead of one per mask).  Both keep the reference's skip of
// |psi_i|^2 < 1e-18, which leaves a small imaginary part on large states.
void iq_statevector_expval_pauli_sum(void* sv, size_t nterms, const double* coeffs, const char* ops, const size_t* qubits,
                                     const size_t* lens, double* out_real, double* out_imag, double* per_term) {
    if (!sv || !ops || !qubits || !lens || !out_real || !out_imag) return;
    Engine* e = sv_of(sv);
    DeviceGuard g(e->dev);
    std::vector<PauliTerm> terms(nterms);
    size_t off = 0;
    for (size_t t = 0; t < nterms; ++t) {
        if (!parse_pauli(ops + off, qubits + off, lens[t], e->n, terms[t])) return;
        off += lens[t];
    }
    if (!per_term) {
        double energy[2] = {0.0, 0.0};
        bool used = false;
        if (!pauli_sum_tiled(*e, nterms, terms.data(), coeffs, energy, &used)) return;
        if (used) {
            *out_real = energy[0];
            *out_imag = energy[1];
            return;
        }
    }
    std::vector<double> vals(2 * std::max<size_t>(nterms, 1), 0.0);
    if (nterms && !e->pauli_terms(nterms, terms.data(), vals.data())) return;
    double sr = 0.0, si = 0.0;
    for (size_t t = 0; t < nterms; ++t) {
        const double c = coeffs ? coeffs[t] : 1.0;
        sr += c * vals[2 * t];
        si += c * vals[2 * t + 1];
    }
    *out_real = sr;
    *out_imag = si;
    if (per_term)
        for (size_t i = 0; i < 2 * nterms; ++i) per_term[i] = vals[i];
}

void iq_statevector_synchronize(void* sv) {
    if (!sv) return;
    DeviceGuard g(sv_of(sv)->dev);
    sv_of(sv)->sync();
}

void* iq_statevector_device_ptr(void* sv) { return sv ? sv_of(sv)->amp : nullptr; }
void* iq_statevector_stream(void* sv) { return sv ? (void*)sv_of(sv)->stream : nullptr; }

// ---------------------------------------------------------- density matrix

void* iq_density_matrix_create(size_t num_qubits) {
    if (num_qubits > 16) {
        set_error(kErrInvalidArgument, "density matrices above 16 qubits are not supported");
        return nullptr;
    }
    DMHandle* h = new (std::nothrow) DMHandle();
    if (!h) return nullptr;
    h->nq = (int)num_qubits;
    if (!h->eng.init(2 * (int)num_qubits, kC128)) {
        delete h;
        return nullptr;
    }
    return h;
}

void iq_density_matrix_free(void* dm) {
    if (!dm) return;
    DeviceGuard g(dm_of(dm)->eng.dev);
    delete dm_of(dm);
}

void iq_density_matrix_reset(void* dm) {
    if (!dm) return;
    DeviceGuard g(dm_of(dm)->eng.dev);
    dm_of(dm)->eng.reset_basis(0);
}

void iq_density_matrix_apply_1q_unitary(void* dm, size_t target, const double* u_real, const double* u_imag) {
    if (!dm || !u_real || !u_imag) return;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    if (target >= (size_t)h->nq) {
        set_error(kErrInvalidArgument, "target qubit %zu out of range", target);
        return;
    }
    const uint32_t q = (uint32_t)target;
    dm_apply_superop(h, &q, 1, 1, u_real, u_imag);
}

void iq_density_matrix_apply_2q_unitary(void* dm, size_t q0, size_t q1, const double* u_real, const double* u_imag) {
    if (!dm || !u_real || !u_imag) return;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    if (q0 >= (size_t)h->nq || q1 >= (size_t)h->nq || q0 == q1) {
        set_error(kErrInvalidArgument, "invalid qubit indices for a 2-qubit unitary");
        return;
    }
    const uint32_t q[2] = {(uint32_t)q0, (uint32_t)q1};
    dm_apply_superop(h, q, 2, 1, u_real, u_imag);
}

void iq_density_matrix_apply_kraus_1q(void* dm, size_t target, size_t num_kraus, const double* k_real, const double* k_imag) {
    if (!dm || !k_real || !k_imag) return;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    if (target >= (size_t)h->nq) {
        set_error(kErrInvalidArgument, "target qubit %zu out of range", target);
        return;
    }
    const uint32_t q = (uint32_t)target;
    dm_apply_superop(h, &q, 1, num_kraus, k_real, k_imag);
}

// Multi-qubit Kraus channel on 1 or 2 qubits (reference
// core/src/density_matrix.cpp:165-191 has it as a class method only).
void iq_density_matrix_apply_kraus(void* dm, const size_t* qubits, size_t k, size_t num_kraus, const double* k_real,
                                   const double* k_imag) {
    if (!dm || !qubits || !k_real || !k_imag) return;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    if (k < 1 || k > 2) {
        set_error(kErrInvalidArgument, "Kraus channels act on 1 or 2 qubits");
        return;
    }
    uint32_t q[2];
    to_u32(qubits, k, q);
    for (size_t i = 0; i < k; ++i)
        if ((int)q[i] >= h->nq || (i == 1 && q[0] == q[1])) {
            set_error(kErrInvalidArgument, "invalid Kraus qubits");
            return;
        }
    dm_apply_superop(h, q, (int)k, num_kraus, k_real, k_imag);
}

// Superoperator composed by the caller (gate and channel fusion) on 1 or 2
// qubits: S is 4^k x 4^k, row-major, on the local index (row bits of
// qubits[0..k-1], then their column bits), MSB-first; rho -> sum K rho K^dagger
// is S = sum K (x) conj(K).  One pass over rho, like a single Kraus channel.
void iq_density_matrix_apply_superop(void* dm, const size_t* qubits, size_t k, const double* s_real,
                                     const double* s_imag) {
    if (!dm || !qubits || !s_real || !s_imag) return;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    if (k < 1 || k > 2) {
        set_error(kErrInvalidArgument, "superoperators act on 1 or 2 qubits");
        return;
    }
    uint32_t q[2];
    to_u32(qubits, k, q);
    for (size_t i = 0; i < k; ++i)
        if ((int)q[i] >= h->nq || (i == 1 && q[0] == q[1])) {
            set_error(kErrInvalidArgument, "invalid superoperator qubits");
            return;
        }
    uint32_t bits[4];
    for (size_t i = 0; i < k; ++i) {
        bits[i] = q[i] + (uint32_t)h->nq;  // row qubits
        bits[k + i] = q[i];                 // column qubits
    }
    h->eng.apply(2 * (int)k, bits, s_real, s_imag, gate_path_override());
}

void iq_density_matrix_get_elements(void* dm, double* real_out, double* imag_out) {
    if (!dm || !real_out || !imag_out) return;
    DeviceGuard g(dm_of(dm)->eng.dev);
    dm_of(dm)->eng.export_split(real_out, imag_out);
}

void iq_density_matrix_set_elements(void* dm, const double* real_in, const double* imag_in) {
    if (!dm || !real_in || !imag_in) return;
    DeviceGuard g(dm_of(dm)->eng.dev);
    dm_of(dm)->eng.import_split(real_in, imag_in);
}

// Interleaved (re, im) FP64 copies of rho, 2 * 4^n doubles, row-major.
void iq_density_matrix_get_elements_interleaved(void* dm, double* out) {
    if (!dm || !out) return;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    h->eng.export_interleaved(out);
}

void iq_density_matrix_set_elements_interleaved(void* dm, const double* in) {
    if (!dm || !in) return;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    h->eng.import_interleaved(in);
}

static bool dm_diag(DMHandle* h, int mode, int target, double* out) {
    const int m = mode == 1 ? h->nq : (mode == 2 ? 2 : 1);
    return hip_ok(reduce_dm_diag(h->eng.amp, h->nq, mode, target, h->eng.red_part, h->eng.red_out, h->eng.stream),
                  "diagonal reduction") &&
           h->eng.fetch(m, out);
}

double iq_density_matrix_trace(void* dm) {
    if (!dm) return 0.0;
    DeviceGuard g(dm_of(dm)->eng.dev);
    double v = 0.0;
    dm_diag(dm_of(dm), 0, 0, &v);
    return v;
}

double iq_density_matrix_purity(void* dm) {
    if (!dm) return 0.0;
    DeviceGuard g(dm_of(dm)->eng.dev);
    double v = 0.0;
    dm_of(dm)->eng.norm2(&v);
    return v;
}

void iq_density_matrix_expval_z(void* dm, double* out) {
    if (!dm || !out) return;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    if (h->nq == 0) return;
    dm_diag(h, 1, 0, out);
}

// reference core/src/density_matrix.cpp:310-380
double iq_density_matrix_entropy(void* dm) {
    if (!dm) return 0.0;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    const size_t d = (size_t)1 << h->nq;
    if (d > 64) {
        double v = 0.0;
        dm_diag(h, 3, 0, &v);
        return v;
    }
    std::vector<double> re(d * d), im(d * d);
    if (!h->eng.export_split(re.data(), im.data())) return std::numeric_limits<double>::quiet_NaN();
    std::vector<std::complex<double>> A(d * d);
    for (size_t i = 0; i < d * d; ++i) A[i] = {re[i], im[i]};
    return entropy_bits(jacobi_eigenvalues(std::move(A), d, 120, 1e-13));
}

// reference core/src/density_matrix.cpp:218-256
int iq_density_matrix_measure(void* dm, size_t target, double rand_val) {
    if (!dm) return 0;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    if (target >= (size_t)h->nq) {
        set_error(kErrInvalidArgument, "target qubit %zu out of range", target);
        return 0;
    }
    double p[2];
    if (!dm_diag(h, 2, (int)target, p)) return 0;
    p[0] = std::max(0.0, p[0]);
    p[1] = std::max(0.0, p[1]);
    const int outcome = (rand_val < p[0]) ? 0 : 1;
    if (p[outcome] < 1e-15) {
        set_error(kErrZeroProbability, "cannot collapse to a zero-probability outcome");
        return outcome;
    }
    hip_ok(launch_dm_collapse(h->eng.amp, h->nq, (int)target, outcome, 1.0 / p[outcome], h->eng.stream), "collapse");
    return outcome;
}

// reference core/src/density_matrix.cpp:258-269: Kraus {|0><0|, |0><1|}.
void iq_density_matrix_reset_qubit(void* dm, size_t target) {
    if (!dm) return;
    DMHandle* h = dm_of(dm);
    DeviceGuard g(h->eng.dev);
    if (target >= (size_t)h->nq) {
        set_error(kErrInvalidArgument, "target qubit %zu out of range", target);
        return;
    }
    const double kr[8] = {1, 0, 0, 0, 0, 1, 0, 0}, ki[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t q = (uint32_t)target;
    dm_apply_superop(h, &q, 1, 2, kr, ki);
}

void iq_density_matrix_synchronize(void* dm) {
    if (!dm) return;
    DeviceGuard g(dm_of(dm)->eng.dev);
    dm_of(dm)->eng.sync();
}

void* iq_density_matrix_stream(void* dm) { return dm ? (void*)dm_of(dm)->eng.stream : nullptr; }

}  // extern "C"
