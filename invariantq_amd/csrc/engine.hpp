// Device-resident state handle shared by the statevector, density-matrix,
// adjoint and batch entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdarg>
#include <vector>

#include "common.hpp"
#include "jacobi.hpp"
#include "kernels.hpp"

namespace iqg {

void set_error(int code, const char* fmt, ...);
bool hip_ok(hipError_t e, const char* what);
int gate_path_override();

class DeviceGuard {
public:
    explicit DeviceGuard(int dev);
    ~DeviceGuard();

private:
    int prev_ = 0, dev_ = 0;
};

struct PauliTerm {
    uint64_t xmask = 0;   // X and Y positions (flip mask)
    uint64_t zymask = 0;  // Y and Z positions (sign mask)
    int ny = 0;           // number of Y factors
};

bool parse_pauli(const char* ops, const size_t* qubits, size_t len, int n, PauliTerm& t);

struct Engine {
    static constexpr int kSlots = 16;
    static constexpr uint64_t kStage = 1ULL << 22;  // export chunk, doubles

    struct Slot {
        void* host = nullptr;
        void* host_dev = nullptr;  // device view of the pinned slot
        void* dev = nullptr;
        hipEvent_t ev = nullptr;
        bool pending = false;
    };

    int n = 0;
    int prec = kC128;
    uint64_t dim = 0;
    size_t elem = 16;
    int dev = 0;
    void* amp = nullptr;
    hipStream_t stream = nullptr;
    double* red_part = nullptr;
    double* red_out = nullptr;
    double* host_out = nullptr;
    std::vector<Slot> slots;
    int next_slot = 0;
    size_t slot_bytes = 0;
    double* work = nullptr;
    size_t work_doubles = 0;
    double* stage = nullptr;
    // host copies of exports and imports: two pinned slots of `arrays` x
    // kStage doubles, so that one chunk's DMA overlaps the previous chunk's
    // host copy (pageable copies through the runtime run at about 16 GB/s)
    double* pinned = nullptr;
    size_t pinned_doubles = 0;
    hipEvent_t pinned_ev[2] = {nullptr, nullptr};

    // matrices of one apply_sequence call: pinned (coherent) + device copy
    struct SeqBuf {
        void* host = nullptr;
        void* host_dev = nullptr;
        void* dev = nullptr;
        size_t cap = 0;
        hipEvent_t ev = nullptr;
        bool pending = false;
    } seq;

    struct GatePlan {
        enum Kind { kDiag, kControlled, kValue, kDevice };
        Kind kind = kValue;
        GateDesc g;
        size_t bytes = 0;
        bool mfma = false;
        bool m3 = false;
        uint32_t ctrl[2] = {}, targ = 0;  // kControlled: U (row-major 2x2) on targ where every ctrl is set
        int nctrl = 0;
        double ure[4] = {}, uim[4] = {};
    };

    Engine() = default;
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    ~Engine();

    bool init(int nqubits, int precision);
    bool reset_basis(uint64_t index);
    const void* upload_matrix(int K, bool m3, const double* re, const double* im, const uint32_t* mfma_q);
    bool check_gate(int K, const uint32_t* q) const;
    GatePlan plan_gate(int K, const uint32_t* q, const double* re, const double* im, int path) const;
    void pack_plan(const GatePlan& p, void* host_dst) const;
    bool launch_plan(GatePlan& p, const void* dmat);
    bool apply(int K, const uint32_t* q, const double* re, const double* im, int path);
    bool apply_sequence(size_t count, const int* ks, const uint32_t* qubits, const double* re, const double* im,
                        int path);
    bool sync();
    bool ensure_work(size_t doubles);
    void trim_work(size_t keep_doubles);  // frees the work buffer when it holds more
    bool ensure_stage();
    bool ensure_pinned(int arrays);
    bool ensure_pinned_doubles(size_t need);
    bool export_chunks(int arrays, double* const* dst, bool (*produce)(Engine*, uint64_t, uint64_t));
    // `count` doubles of device memory into pageable host memory through the
    // pinned slots (large copies), waiting for the handle's stream first
    bool download(const double* dsrc, double* hdst, uint64_t count);
    bool export_split(double* re, double* im);
    bool import_split(const double* re, const double* im);
    // interleaved FP64 (re, im) pairs, 2 * dim doubles
    bool export_interleaved(double* out);
    bool import_interleaved(const double* in);
    bool export_probs(double* out);
    bool fetch(int m, double* out);
    bool norm2(double* out);
    bool bit_probs(int target, double* p01);
    bool pauli_terms(size_t nterms, const PauliTerm* terms, double* vals);
};

// Sum_t coeff_t <psi|P_t|psi> of the engine's state over tile passes
// (adjoint.hip), Re and Im into out[0], out[1]; *used = false when one pass
// per flip mask is as cheap.
bool pauli_sum_tiled(Engine& e, size_t nterms, const PauliTerm* terms, const double* coeffs, double* out, bool* used);

}  // namespace iqg
