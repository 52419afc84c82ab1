// Parametric circuits on the GPU: forward evaluation of <H>, adjoint-method
// gradients and (for comparison) parameter-shift gradients.
//
// No reference counterpart: InvariantQ has no gradient code (its VQE,
// invariantq/benchmarks/vqe.py:44-98, is gradient-free COBYLA over
// StateVectorSimulator.run).  Gate matrices follow invariantq/ir/qir.py:90-136
// and Pauli strings follow core/src/statevector.cpp:258-291.
//
// Adjoint method, for U_k = exp(-i theta_k G_k / 2) with Pauli generator G_k:
//   psi_L = U_L..U_1 |0>,  lambda = H psi_L,
//   for k = L..1:  dE/dtheta_k = Im <lambda | G_k psi_k>,  psi <- U_k^dag psi,
//                  lambda <- U_k^dag lambda.
// U3(theta, phi, lam) is taken as RZ(phi) RY(theta) RZ(lam) (equal up to a
// global phase, which leaves <H> and its gradient unchanged).
//
// The tape is compiled into segments, reordering only commuting gates:
//   1q   a run of one-qubit gates on one qubit (<= 4 sub-steps): forward is one
//        pass with the product matrix; reverse is one fused pass that walks the
//        sub-steps backwards, accumulating each generator inner product in
//        FP64 and un-applying the gate to psi and lambda in registers;
//   diag a set of CZ gates (mutually commuting): one sign pass;
//   2q   any other two-qubit gate.
// Each reverse pass moves 4 * 2^n * sizeof(amplitude) bytes, the algorithmic
// minimum for updating two states.  Per-block partial sums of every segment
// are reduced by one final kernel, so the sweep never synchronises with the host.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <new>
#include <vector>

#include "engine.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

namespace iqg {

enum GateKind : int { kRX = 0, kRY = 1, kRZ = 2, kU3 = 3, kCZ = 4, kCX = 5, kMat1 = 6, kMat2 = 7 };
enum Gen : int { kGenNone = 0, kGenX = 1, kGenY = 2, kGenZ = 3 };

constexpr int kSubMax = 4;
constexpr int kEdgeMax = 64;

namespace {

struct Seg1qArgs {
    double u[kSubMax][8];  // U^dagger of each sub-step: re00 im00 re01 im01 re10 im10 re11 im11
    int gen[kSubMax];
    int count;
};

struct EdgeList {
    uint8_t a[kEdgeMax];
    uint8_t b[kEdgeMax];
    int ne;
};

struct ObsTable {
    const uint64_t* xmask;  // per group
    const int* start;       // group g owns terms [start[g], start[g+1])
    const uint64_t* zy;     // per term
    const int* ny;
    const double* coeff;
    int ngroups;
};

template <typename R>
__device__ __forceinline__ void apply2(const double* u, double& x0r, double& x0i, double& x1r, double& x1i) {
    const double y0r = u[0] * x0r - u[1] * x0i + u[2] * x1r - u[3] * x1i;
    const double y0i = u[0] * x0i + u[1] * x0r + u[2] * x1i + u[3] * x1r;
    const double y1r = u[4] * x0r - u[5] * x0i + u[6] * x1r - u[7] * x1i;
    const double y1i = u[4] * x0i + u[5] * x0r + u[6] * x1i + u[7] * x1r;
    x0r = (double)(R)y0r;
    x0i = (double)(R)y0i;
    x1r = (double)(R)y1r;
    x1i = (double)(R)y1i;
}

// Reverse pass over a run of one-qubit gates on qubit q.  Amplitudes are
// widened to FP64 in registers; each sub-step's result is rounded back to the
// state precision, so the arithmetic matches un-applying the gates one by one.
template <typename R>
__global__ void __launch_bounds__(256) k_adj_seg1q(cv<R>* __restrict__ psi, cv<R>* __restrict__ lam, uint64_t npairs,
                                                   int q, Seg1qArgs a, double* __restrict__ part) {
    double acc[kSubMax] = {0.0, 0.0, 0.0, 0.0};
    const uint64_t bit = 1ULL << q, lo = bit - 1;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += stride) {
        const uint64_t i0 = ((p & ~lo) << 1) | (p & lo), i1 = i0 | bit;
        const cv<R> A0 = psi[i0], A1 = psi[i1], B0 = lam[i0], B1 = lam[i1];
        double a0r = A0.x, a0i = A0.y, a1r = A1.x, a1i = A1.y;
        double b0r = B0.x, b0i = B0.y, b1r = B1.x, b1i = B1.y;
#pragma unroll
        for (int s = kSubMax - 1; s >= 0; --s) {
            if (s >= a.count) continue;
            const int g = a.gen[s];
            if (g == kGenX) {
                acc[s] += (b0r * a1i - b0i * a1r) + (b1r * a0i - b1i * a0r);
            } else if (g == kGenY) {
                acc[s] += (-b0r * a1r - b0i * a1i) + (b1r * a0r + b1i * a0i);
            } else if (g == kGenZ) {
                acc[s] += (b0r * a0i - b0i * a0r) - (b1r * a1i - b1i * a1r);
            }
            apply2<R>(a.u[s], a0r, a0i, a1r, a1i);
            apply2<R>(a.u[s], b0r, b0i, b1r, b1i);
        }
        cv<R> o;
        o.x = (R)a0r; o.y = (R)a0i; psi[i0] = o;
        o.x = (R)a1r; o.y = (R)a1i; psi[i1] = o;
        o.x = (R)b0r; o.y = (R)b0i; lam[i0] = o;
        o.x = (R)b1r; o.y = (R)b1i; lam[i1] = o;
    }
    block_sum_store<kSubMax>(acc, kSubMax, part);
}

// Product of CZ signs over an edge list, on psi and (optionally) lambda.
template <typename R>
__global__ void __launch_bounds__(256) k_cz_set(cv<R>* __restrict__ psi, cv<R>* __restrict__ lam, uint64_t count, EdgeList e) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        uint32_t par = 0;
        for (int k = 0; k < e.ne; ++k) par ^= (uint32_t)((i >> e.a[k]) & (i >> e.b[k]) & 1ULL);
        if (par) {
            cv<R> v = psi[i];
            v.x = -v.x;
            v.y = -v.y;
            psi[i] = v;
            if (lam) {
                cv<R> w = lam[i];
                w.x = -w.x;
                w.y = -w.y;
                lam[i] = w;
            }
        }
    }
}

// lambda = H psi and E = Re <psi|H psi> in one pass:
// (H psi)[k] = sum_t c_t phase_t(k ^ x_t) psi[k ^ x_t], phase(j) = i^{nY} (-1)^{popcount(j & (Y|Z))}.
template <typename R, bool WRITE>
__global__ void __launch_bounds__(256) k_apply_h(const cv<R>* __restrict__ psi, cv<R>* __restrict__ lam, uint64_t count,
                                                 ObsTable h, double* __restrict__ part) {
    double e[1] = {0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
        double sr = 0.0, si = 0.0;
        for (int g = 0; g < h.ngroups; ++g) {
            const uint64_t j = k ^ h.xmask[g];
            const cv<R> v = psi[j];
            const double vr = v.x, vi = v.y;
            double cr = 0.0, ci = 0.0;  // sum of c_t * phase_t(j)
            for (int t = h.start[g]; t < h.start[g + 1]; ++t) {
                const double sg = (__popcll(j & h.zy[t]) & 1) ? -h.coeff[t] : h.coeff[t];
                switch (h.ny[t] & 3) {
                    case 0: cr += sg; break;
                    case 1: ci += sg; break;
                    case 2: cr -= sg; break;
                    default: ci -= sg; break;
                }
            }
            sr += cr * vr - ci * vi;
            si += cr * vi + ci * vr;
        }
        const cv<R> p = psi[k];
        e[0] += (double)p.x * sr + (double)p.y * si;
        if (WRITE) {
            cv<R> o;
            o.x = (R)sr;
            o.y = (R)si;
            lam[k] = o;
        }
    }
    block_sum_store<1>(e, 1, part);
}

__global__ void __launch_bounds__(256) k_finalize_slots(const double* __restrict__ part, int nblocks, int nslots,
                                                        double* __restrict__ out) {
    // one block per output: slot = blockIdx.x / kSubMax, sub-step = blockIdx.x % kSubMax
    const int slot = blockIdx.x / kSubMax, s = blockIdx.x % kSubMax;
    if (slot >= nslots) return;
    __shared__ double sh[4];
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) v += part[((uint64_t)slot * nblocks + b) * kSubMax + s];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

inline unsigned nblocks_for(uint64_t items) {
    uint64_t g = (items + kRedThreads - 1) / kRedThreads;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, kRedBlocks));
}

using cplx = std::complex<double>;
struct M2 {
    cplx m[4];
};

M2 mat_rx(double t) { const double c = std::cos(t / 2), s = std::sin(t / 2); return {{c, cplx(0, -s), cplx(0, -s), c}}; }
M2 mat_ry(double t) { const double c = std::cos(t / 2), s = std::sin(t / 2); return {{c, -s, s, c}}; }
M2 mat_rz(double t) { return {{std::exp(cplx(0, -t / 2)), 0.0, 0.0, std::exp(cplx(0, t / 2))}}; }
M2 mul(const M2& a, const M2& b) {  // a * b
    return {{a.m[0] * b.m[0] + a.m[1] * b.m[2], a.m[0] * b.m[1] + a.m[1] * b.m[3],
             a.m[2] * b.m[0] + a.m[3] * b.m[2], a.m[2] * b.m[1] + a.m[3] * b.m[3]}};
}
M2 dag(const M2& a) { return {{std::conj(a.m[0]), std::conj(a.m[2]), std::conj(a.m[1]), std::conj(a.m[3])}}; }

}  // namespace

struct TapeGate {
    int kind = 0;
    int q0 = 0, q1 = 0;
    int64_t p[3] = {-1, -1, -1};
    double mre[16] = {}, mim[16] = {};
};

struct Sub {
    int gate;
    int which;  // U3 sub-step: 0 = RZ(lam), 1 = RY(theta), 2 = RZ(phi)
};

struct Segment {
    int kind = 0;  // 0 = 1q, 1 = diag, 2 = 2q
    int q = -1;
    std::vector<Sub> subs;
    std::vector<std::pair<int, int>> edges;
    int gate = -1;
    int slot = -1;
};

struct Circuit {
    int n = 0, prec = kC128;
    Engine eng;
    void* lam = nullptr;
    std::vector<TapeGate> gates;
    std::vector<Segment> segs;
    bool compiled = false;
    int nslots = 0;
    unsigned nblk = 1;
    double* gpart = nullptr;
    double* gout = nullptr;
    double* hgout = nullptr;
    size_t gcap = 0;
    // observable
    std::vector<double> coeff;
    std::vector<PauliTerm> terms;
    void* obs_dev = nullptr;
    ObsTable obs{};
    double* evals = nullptr;  // device energies for parameter shift
    size_t evals_cap = 0;

    ~Circuit() {
        if (eng.stream) hipStreamSynchronize(eng.stream);
        if (lam) hipFree(lam);
        if (gpart) hipFree(gpart);
        if (gout) hipFree(gout);
        if (hgout) hipHostFree(hgout);
        if (obs_dev) hipFree(obs_dev);
        if (evals) hipFree(evals);
    }
};

namespace {

M2 sub_matrix(const TapeGate& g, int which, const double* params) {
    auto P = [&](int i) { return params[g.p[i]]; };
    switch (g.kind) {
        case kRX: return mat_rx(P(0));
        case kRY: return mat_ry(P(0));
        case kRZ: return mat_rz(P(0));
        case kU3: return which == 0 ? mat_rz(P(2)) : (which == 1 ? mat_ry(P(0)) : mat_rz(P(1)));
        default: {
            M2 m;
            for (int i = 0; i < 4; ++i) m.m[i] = cplx(g.mre[i], g.mim[i]);
            return m;
        }
    }
}

int sub_gen(const TapeGate& g, int which) {
    switch (g.kind) {
        case kRX: return kGenX;
        case kRY: return kGenY;
        case kRZ: return kGenZ;
        case kU3: return which == 1 ? kGenY : kGenZ;
        default: return kGenNone;
    }
}

int64_t sub_param(const TapeGate& g, int which) {
    switch (g.kind) {
        case kRX: case kRY: case kRZ: return g.p[0];
        case kU3: return which == 0 ? g.p[2] : (which == 1 ? g.p[0] : g.p[1]);
        default: return -1;
    }
}

bool is_1q(int kind) { return kind == kRX || kind == kRY || kind == kRZ || kind == kU3 || kind == kMat1; }

// Dependency-respecting reordering: ops sharing a qubit keep their order,
// except CZ-CZ pairs, which commute.  Greedy list scheduling that emits runs of
// one-qubit gates per qubit, then every ready CZ as one sign pass.
void compile(Circuit& c) {
    const int L = (int)c.gates.size();
    std::vector<std::vector<int>> preds(L), succs(L);
    std::vector<int> last_nondiag(c.n, -1);
    std::vector<std::vector<int>> diag_since(c.n);
    for (int i = 0; i < L; ++i) {
        const TapeGate& g = c.gates[i];
        const bool diag = g.kind == kCZ;
        std::vector<int> qs = {g.q0};
        if (!is_1q(g.kind)) qs.push_back(g.q1);
        for (int q : qs) {
            if (last_nondiag[q] >= 0) preds[i].push_back(last_nondiag[q]);
            if (diag) {
                diag_since[q].push_back(i);
            } else {
                for (int d : diag_since[q]) preds[i].push_back(d);
                diag_since[q].clear();
                last_nondiag[q] = i;
            }
        }
        std::sort(preds[i].begin(), preds[i].end());
        preds[i].erase(std::unique(preds[i].begin(), preds[i].end()), preds[i].end());
        for (int p : preds[i]) succs[p].push_back(i);
    }
    std::vector<int> indeg(L), done(L, 0);
    for (int i = 0; i < L; ++i) indeg[i] = (int)preds[i].size();
    auto ready = [&](int i) { return !done[i] && indeg[i] == 0; };
    auto finish = [&](int i) {
        done[i] = 1;
        for (int s : succs[i]) --indeg[s];
    };
    c.segs.clear();
    int emitted = 0;
    while (emitted < L) {
        bool progress = false;
        // runs of one-qubit gates, qubit by qubit
        for (int q = 0; q < c.n; ++q) {
            Segment seg;
            seg.kind = 0;
            seg.q = q;
            for (;;) {
                int pick = -1;
                for (int i = 0; i < L; ++i)
                    if (ready(i) && is_1q(c.gates[i].kind) && c.gates[i].q0 == q) { pick = i; break; }
                if (pick < 0) break;
                const int nsub = c.gates[pick].kind == kU3 ? 3 : 1;
                if ((int)seg.subs.size() + nsub > kSubMax) {
                    c.segs.push_back(seg);
                    seg.subs.clear();
                }
                for (int w = 0; w < nsub; ++w) seg.subs.push_back({pick, w});
                finish(pick);
                ++emitted;
            }
            if (!seg.subs.empty()) {
                c.segs.push_back(seg);
                progress = true;
            }
        }
        // every ready CZ as one sign pass
        Segment d;
        d.kind = 1;
        for (int i = 0; i < L; ++i)
            if (ready(i) && c.gates[i].kind == kCZ) {
                if ((int)d.edges.size() == kEdgeMax) {
                    c.segs.push_back(d);
                    d.edges.clear();
                }
                d.edges.push_back({c.gates[i].q0, c.gates[i].q1});
                finish(i);
                ++emitted;
            }
        if (!d.edges.empty()) {
            c.segs.push_back(d);
            progress = true;
        }
        if (progress) continue;
        for (int i = 0; i < L; ++i)
            if (ready(i)) {
                Segment s;
                s.kind = 2;
                s.gate = i;
                c.segs.push_back(s);
                finish(i);
                ++emitted;
                break;
            }
    }
    c.nslots = 0;
    for (auto& s : c.segs)
        if (s.kind == 0) s.slot = c.nslots++;
    c.compiled = true;
}

bool ensure_grad_buffers(Circuit& c) {
    const size_t need = (size_t)std::max(c.nslots, 1);
    if (c.gcap >= need) return true;
    if (c.gpart) hipFree(c.gpart);
    if (c.gout) hipFree(c.gout);
    if (c.hgout) hipHostFree(c.hgout);
    c.gpart = c.gout = c.hgout = nullptr;
    c.gcap = 0;
    if (!hip_ok(hipMalloc(&c.gpart, need * c.nblk * kSubMax * sizeof(double)), "gradient partials") ||
        !hip_ok(hipMalloc(&c.gout, need * kSubMax * sizeof(double)), "gradient slots") ||
        !hip_ok(hipHostMalloc((void**)&c.hgout, need * kSubMax * sizeof(double)), "gradient host slots"))
        return false;
    c.gcap = need;
    return true;
}

bool check_params(const Circuit& c, size_t nparams) {
    for (const auto& g : c.gates)
        for (int i = 0; i < 3; ++i)
            if (g.p[i] >= 0 && (size_t)g.p[i] >= nparams) {
                set_error(kErrInvalidArgument, "gate uses parameter %lld but only %zu were given", (long long)g.p[i], nparams);
                return false;
            }
    return true;
}

template <typename R>
bool run_forward(Circuit& c, const double* params) {
    Engine& e = c.eng;
    if (!e.reset_basis(0)) return false;
    for (const auto& s : c.segs) {
        if (s.kind == 0) {
            M2 m = {{1.0, 0.0, 0.0, 1.0}};
            for (const auto& sub : s.subs) m = mul(sub_matrix(c.gates[sub.gate], sub.which, params), m);
            double re[4], im[4];
            for (int i = 0; i < 4; ++i) {
                re[i] = m.m[i].real();
                im[i] = m.m[i].imag();
            }
            GateDesc g;
            g.K = 1;
            g.q[0] = (uint32_t)s.q;
            g.re = re;
            g.im = im;
            if (!hip_ok(launch_gate(e.amp, c.n, c.prec, e.stream, g), "forward 1q")) return false;
        } else if (s.kind == 1) {
            EdgeList el{};
            el.ne = (int)s.edges.size();
            for (int k = 0; k < el.ne; ++k) {
                el.a[k] = (uint8_t)s.edges[k].first;
                el.b[k] = (uint8_t)s.edges[k].second;
            }
            const uint64_t count = e.dim;
            const unsigned b = (unsigned)std::min<uint64_t>((count + 255) / 256, 65536);
            hipLaunchKernelGGL(k_cz_set<R>, dim3(b), dim3(256), 0, e.stream, (cv<R>*)e.amp, (cv<R>*)nullptr, count, el);
            if (!hip_ok(hipGetLastError(), "forward CZ")) return false;
        } else {
            const TapeGate& tg = c.gates[s.gate];
            double re[16], im[16];
            if (tg.kind == kCX) {
                const double cx[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0};
                std::memcpy(re, cx, sizeof(re));
                std::memset(im, 0, sizeof(im));
            } else {
                std::memcpy(re, tg.mre, sizeof(re));
                std::memcpy(im, tg.mim, sizeof(im));
            }
            GateDesc g;
            g.K = 2;
            g.q[0] = (uint32_t)tg.q0;
            g.q[1] = (uint32_t)tg.q1;
            g.re = re;
            g.im = im;
            if (!hip_ok(launch_gate(e.amp, c.n, c.prec, e.stream, g), "forward 2q")) return false;
        }
    }
    return true;
}

template <typename R>
bool apply_observable(Circuit& c, bool write, double* energy_dev) {
    Engine& e = c.eng;
    const unsigned b = nblocks_for(e.dim);
    if (write)
        hipLaunchKernelGGL((k_apply_h<R, true>), dim3(b), dim3(256), 0, e.stream, (const cv<R>*)e.amp, (cv<R>*)c.lam, e.dim,
                           c.obs, e.red_part);
    else
        hipLaunchKernelGGL((k_apply_h<R, false>), dim3(b), dim3(256), 0, e.stream, (const cv<R>*)e.amp, (cv<R>*)nullptr,
                           e.dim, c.obs, e.red_part);
    return hip_ok(hipGetLastError(), "observable") && hip_ok(launch_finalize(e.red_part, b, 1, energy_dev, e.stream), "finalize");
}

template <typename R>
bool run_reverse(Circuit& c, const double* params) {
    Engine& e = c.eng;
    const uint64_t npairs = e.dim >> 1;
    for (int si = (int)c.segs.size() - 1; si >= 0; --si) {
        const Segment& s = c.segs[si];
        if (s.kind == 0) {
            Seg1qArgs a{};
            a.count = (int)s.subs.size();
            for (int k = 0; k < a.count; ++k) {
                const TapeGate& g = c.gates[s.subs[k].gate];
                const M2 d = dag(sub_matrix(g, s.subs[k].which, params));
                for (int i = 0; i < 4; ++i) {
                    a.u[k][2 * i] = d.m[i].real();
                    a.u[k][2 * i + 1] = d.m[i].imag();
                }
                a.gen[k] = sub_gen(g, s.subs[k].which);
            }
            hipLaunchKernelGGL(k_adj_seg1q<R>, dim3(c.nblk), dim3(256), 0, e.stream, (cv<R>*)e.amp, (cv<R>*)c.lam, npairs,
                               s.q, a, c.gpart + (size_t)s.slot * c.nblk * kSubMax);
            if (!hip_ok(hipGetLastError(), "reverse 1q")) return false;
        } else if (s.kind == 1) {
            EdgeList el{};
            el.ne = (int)s.edges.size();
            for (int k = 0; k < el.ne; ++k) {
                el.a[k] = (uint8_t)s.edges[k].first;
                el.b[k] = (uint8_t)s.edges[k].second;
            }
            const unsigned b = (unsigned)std::min<uint64_t>((e.dim + 255) / 256, 65536);
            hipLaunchKernelGGL(k_cz_set<R>, dim3(b), dim3(256), 0, e.stream, (cv<R>*)e.amp, (cv<R>*)c.lam, e.dim, el);
            if (!hip_ok(hipGetLastError(), "reverse CZ")) return false;
        } else {
            const TapeGate& tg = c.gates[s.gate];
            double re[16], im[16];
            if (tg.kind == kCX) {  // CX is its own inverse
                const double cx[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0};
                std::memcpy(re, cx, sizeof(re));
                std::memset(im, 0, sizeof(im));
            } else {
                for (int r = 0; r < 4; ++r)
                    for (int col = 0; col < 4; ++col) {
                        re[r * 4 + col] = tg.mre[col * 4 + r];
                        im[r * 4 + col] = -tg.mim[col * 4 + r];
                    }
            }
            GateDesc g;
            g.K = 2;
            g.q[0] = (uint32_t)tg.q0;
            g.q[1] = (uint32_t)tg.q1;
            g.re = re;
            g.im = im;
            if (!hip_ok(launch_gate(e.amp, c.n, c.prec, e.stream, g), "reverse 2q psi") ||
                !hip_ok(launch_gate(c.lam, c.n, c.prec, e.stream, g), "reverse 2q lambda"))
                return false;
        }
    }
    if (c.nslots > 0) {
        hipLaunchKernelGGL(k_finalize_slots, dim3(c.nslots * kSubMax), dim3(256), 0, e.stream, c.gpart, (int)c.nblk,
                           c.nslots, c.gout);
        if (!hip_ok(hipGetLastError(), "gradient finalize")) return false;
    }
    return true;
}

Circuit* circ_of(void* p) { return static_cast<Circuit*>(p); }

bool prepare(Circuit& c, const double* params, size_t nparams) {
    if (c.terms.empty()) {
        set_error(kErrInvalidArgument, "no observable set");
        return false;
    }
    if (!check_params(c, nparams)) return false;
    if (!c.compiled) compile(c);
    return true;
}

}  // namespace
}  // namespace iqg

using namespace iqg;

extern "C" {

void* iq_circuit_create(size_t num_qubits, int precision) {
    if (num_qubits < 1 || num_qubits > 34 || (precision != kC128 && precision != kC64)) {
        set_error(kErrInvalidArgument, "unsupported circuit shape (n=%zu, precision=%d)", num_qubits, precision);
        return nullptr;
    }
    Circuit* c = new (std::nothrow) Circuit();
    if (!c) return nullptr;
    c->n = (int)num_qubits;
    c->prec = precision;
    if (!c->eng.init((int)num_qubits, precision) || hipMalloc(&c->lam, c->eng.dim * c->eng.elem) != hipSuccess) {
        if (!c->lam) set_error(kErrOutOfMemory, "cannot allocate the adjoint state");
        delete c;
        return nullptr;
    }
    c->nblk = nblocks_for(c->eng.dim >> 1);
    return c;
}

void iq_circuit_free(void* c) {
    if (!c) return;
    DeviceGuard g(circ_of(c)->eng.dev);
    delete circ_of(c);
}

// kind: 0 RX, 1 RY, 2 RZ (param_index[0]); 3 U3 (param_index[0..2] = theta, phi, lambda);
// 4 CZ, 5 CX (q0 control); 6 fixed 2x2 matrix; 7 fixed 4x4 matrix (MSB-first on q0, q1).
int iq_circuit_add_gate(void* cp, int kind, size_t q0, size_t q1, const int64_t* param_index, const double* m_re,
                        const double* m_im) {
    if (!cp) return -1;
    Circuit& c = *circ_of(cp);
    TapeGate g;
    g.kind = kind;
    g.q0 = (int)std::min<size_t>(q0, 1 << 20);
    g.q1 = (int)std::min<size_t>(q1, 1 << 20);
    const bool two = (kind == kCZ || kind == kCX || kind == kMat2);
    if (kind < 0 || kind > kMat2 || g.q0 >= c.n || (two && (g.q1 >= c.n || g.q1 == g.q0))) {
        set_error(kErrInvalidArgument, "invalid gate (kind %d, qubits %zu, %zu)", kind, q0, q1);
        return -1;
    }
    const int np = (kind == kU3) ? 3 : (kind <= kRZ ? 1 : 0);
    for (int i = 0; i < np; ++i) {
        if (!param_index || param_index[i] < 0) {
            set_error(kErrInvalidArgument, "parametric gate needs non-negative parameter indices");
            return -1;
        }
        g.p[i] = param_index[i];
    }
    if (kind == kMat1 || kind == kMat2) {
        if (!m_re || !m_im) {
            set_error(kErrInvalidArgument, "fixed-matrix gate needs a matrix");
            return -1;
        }
        const int d = kind == kMat1 ? 4 : 16;
        for (int i = 0; i < d; ++i) {
            g.mre[i] = m_re[i];
            g.mim[i] = m_im[i];
        }
    }
    c.gates.push_back(g);
    c.compiled = false;
    return (int)c.gates.size() - 1;
}

size_t iq_circuit_num_gates(void* c) { return c ? circ_of(c)->gates.size() : 0; }

// Number of HBM passes of the compiled tape: forward segments (the reverse
// sweep has the same count, each pass touching both states).
size_t iq_circuit_num_segments(void* cp) {
    if (!cp) return 0;
    Circuit& c = *circ_of(cp);
    if (!c.compiled) compile(c);
    return c.segs.size();
}

int iq_circuit_set_observable(void* cp, size_t nterms, const double* coeffs, const char* ops, const size_t* qubits,
                              const size_t* lens) {
    if (!cp || !ops || !qubits || !lens || nterms == 0) return -1;
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    std::vector<PauliTerm> terms(nterms);
    size_t off = 0;
    for (size_t t = 0; t < nterms; ++t) {
        if (!parse_pauli(ops + off, qubits + off, lens[t], c.n, terms[t])) return -1;
        off += lens[t];
    }
    // group by flip mask
    std::vector<size_t> order(nterms);
    for (size_t i = 0; i < nterms; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return terms[a].xmask < terms[b].xmask; });
    std::vector<uint64_t> xm, zy;
    std::vector<int> start, ny;
    std::vector<double> co;
    for (size_t k = 0; k < nterms; ++k) {
        const PauliTerm& t = terms[order[k]];
        if (xm.empty() || xm.back() != t.xmask) {
            xm.push_back(t.xmask);
            start.push_back((int)zy.size());
        }
        zy.push_back(t.zymask);
        ny.push_back(t.ny);
        co.push_back(coeffs ? coeffs[order[k]] : 1.0);
    }
    start.push_back((int)zy.size());
    const size_t G = xm.size(), T = zy.size();
    const size_t bytes = G * 8 + (G + 1) * 4 + T * 8 + T * 4 + T * 8 + 64;
    if (c.obs_dev) hipFree(c.obs_dev);
    c.obs_dev = nullptr;
    if (!hip_ok(hipMalloc(&c.obs_dev, bytes), "observable table")) return -1;
    std::vector<char> host(bytes, 0);
    char* base = host.data();
    size_t o = 0;
    auto put = [&](const void* src, size_t n) {
        std::memcpy(base + o, src, n);
        const size_t at = o;
        o += (n + 7) & ~size_t(7);
        return at;
    };
    const size_t oxm = put(xm.data(), G * 8), ost = put(start.data(), (G + 1) * 4), ozy = put(zy.data(), T * 8),
                 ony = put(ny.data(), T * 4), oco = put(co.data(), T * 8);
    if (!hip_ok(hipMemcpy(c.obs_dev, host.data(), bytes, hipMemcpyHostToDevice), "observable upload")) return -1;
    char* d = static_cast<char*>(c.obs_dev);
    c.obs.xmask = reinterpret_cast<const uint64_t*>(d + oxm);
    c.obs.start = reinterpret_cast<const int*>(d + ost);
    c.obs.zy = reinterpret_cast<const uint64_t*>(d + ozy);
    c.obs.ny = reinterpret_cast<const int*>(d + ony);
    c.obs.coeff = reinterpret_cast<const double*>(d + oco);
    c.obs.ngroups = (int)G;
    c.terms = terms;
    return 0;
}

double iq_circuit_expectation(void* cp, const double* params, size_t nparams) {
    if (!cp) return 0.0;
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    if (!prepare(c, params, nparams)) return 0.0;
    const bool ok = c.prec == kC128 ? (run_forward<double>(c, params) && apply_observable<double>(c, false, c.eng.red_out))
                                    : (run_forward<float>(c, params) && apply_observable<float>(c, false, c.eng.red_out));
    double e = 0.0;
    if (ok) c.eng.fetch(1, &e);
    return e;
}

// Adjoint gradient: grad_out[j] = d<H>/d params[j] (summed over every gate using j).
int iq_circuit_gradient(void* cp, const double* params, size_t nparams, double* energy_out, double* grad_out) {
    if (!cp || (!params && nparams) || !grad_out) return -1;
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    if (!prepare(c, params, nparams) || !ensure_grad_buffers(c)) return -1;
    const bool ok = c.prec == kC128
                        ? (run_forward<double>(c, params) && apply_observable<double>(c, true, c.eng.red_out) &&
                           run_reverse<double>(c, params))
                        : (run_forward<float>(c, params) && apply_observable<float>(c, true, c.eng.red_out) &&
                           run_reverse<float>(c, params));
    if (!ok) return -1;
    double e = 0.0;
    if (c.nslots > 0 &&
        !hip_ok(hipMemcpyAsync(c.hgout, c.gout, (size_t)c.nslots * kSubMax * 8, hipMemcpyDeviceToHost, c.eng.stream), "D2H"))
        return -1;
    if (!c.eng.fetch(1, &e)) return -1;
    for (size_t j = 0; j < nparams; ++j) grad_out[j] = 0.0;
    for (const auto& s : c.segs) {
        if (s.kind != 0) continue;
        for (size_t k = 0; k < s.subs.size(); ++k) {
            const int64_t p = sub_param(c.gates[s.subs[k].gate], s.subs[k].which);
            if (p >= 0) grad_out[p] += c.hgout[(size_t)s.slot * kSubMax + k];
        }
    }
    if (energy_out) *energy_out = e;
    return 0;
}

// Parameter-shift gradient on the same engine, for the parameters listed in
// `which` (all when which == NULL): (E(theta + pi/2) - E(theta - pi/2)) / 2,
// exact when each listed parameter drives one Pauli rotation.
int iq_circuit_parameter_shift(void* cp, const double* params, size_t nparams, const int64_t* which, size_t nwhich,
                               double* grad_out) {
    if (!cp || !params || !grad_out) return -1;
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    if (!prepare(c, params, nparams)) return -1;
    const size_t count = which ? nwhich : nparams;
    if (c.evals_cap < 2 * count) {
        if (c.evals) hipFree(c.evals);
        c.evals = nullptr;
        c.evals_cap = 0;
        if (!hip_ok(hipMalloc(&c.evals, std::max<size_t>(2 * count, 1) * 8), "shift energies")) return -1;
        c.evals_cap = 2 * count;
    }
    std::vector<double> shifted(params, params + nparams);
    const double h = M_PI / 2;
    for (size_t k = 0; k < count; ++k) {
        const int64_t j = which ? which[k] : (int64_t)k;
        if (j < 0 || (size_t)j >= nparams) {
            set_error(kErrInvalidArgument, "parameter index out of range");
            return -1;
        }
        for (int sgn = 0; sgn < 2; ++sgn) {
            shifted[j] = params[j] + (sgn == 0 ? h : -h);
            const bool ok = c.prec == kC128
                                ? (run_forward<double>(c, shifted.data()) && apply_observable<double>(c, false, c.evals + 2 * k + sgn))
                                : (run_forward<float>(c, shifted.data()) && apply_observable<float>(c, false, c.evals + 2 * k + sgn));
            if (!ok) return -1;
        }
        shifted[j] = params[j];
    }
    std::vector<double> ev(2 * count);
    if (count && (!hip_ok(hipMemcpyAsync(ev.data(), c.evals, 2 * count * 8, hipMemcpyDeviceToHost, c.eng.stream), "D2H") ||
                  !c.eng.sync()))
        return -1;
    for (size_t k = 0; k < count; ++k) grad_out[k] = 0.5 * (ev[2 * k] - ev[2 * k + 1]);
    return 0;
}

// Amplitudes after the most recent forward pass (or adjoint sweep: the reverse
// sweep returns psi to |0..0> up to rounding).
void iq_circuit_get_amplitudes(void* cp, double* re, double* im) {
    if (!cp || !re || !im) return;
    DeviceGuard dg(circ_of(cp)->eng.dev);
    circ_of(cp)->eng.export_split(re, im);
}

void* iq_circuit_stream(void* cp) { return cp ? (void*)circ_of(cp)->eng.stream : nullptr; }

}  // extern "C"
