#pragma once
// Shared internals of the engine's translation units: error handling, the
// observable program reader, the state handle and its scratch.
//
// Host orchestration: every entry point validates its arguments, lowers the
// request (lower.hpp) and enqueues kernels (kernels_*.hpp) on the handle's
// HIP stream.  There is no CPU compute path: if the device work cannot be
// enqueued the call fails with PL_ERR_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <limits>
#include <thread>
#include <vector>

#include "../../include/pl_amd.h"
#include "kernels_gate.hpp"
#include "lower.hpp"

namespace plamd {

inline thread_local std::string g_err;
// bytes moved by pl_sv_upload / pl_sv_download since the library was loaded
inline std::atomic<uint64_t> g_h2d_bytes{0}, g_d2h_bytes{0};

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIP_OK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            throw DeviceError(std::string(#expr) + ": " + hipGetErrorString(e_));          \
        }                                                                                  \
    } while (0)

inline int device_cus() {
    static int cus = [] {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
            return 256;
        }
        return p.multiProcessorCount;
    }();
    return cus;
}

// Grid for a grid-stride streaming kernel over `work` items, `per` items per
// thread per iteration: enough workgroups to fill every CU several times,
// but no more than the work needs.
inline unsigned grid_for(uint64_t work, int per = 1, unsigned cap_per_cu = 32) {
    const uint64_t need = (work + uint64_t(kBlock) * per - 1) / (uint64_t(kBlock) * per);
    const uint64_t cap = uint64_t(device_cus()) * cap_per_cu;
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min(need, cap)));
}

// ---------------------------------------------------------------------------
// Observable tree parsed from the program encoding of pl_amd.h
// ---------------------------------------------------------------------------
struct ObsNode {
    int kind = 0; // 0 named, 1 hermitian, 2 tensor, 3 hamiltonian
    int op = 0;
    std::vector<int> wires;
    std::vector<std::complex<double>> mat;
    std::vector<double> coeffs;
    std::vector<ObsNode> kids;
};

struct ObsReader {
    const int *prog;
    const double *coeffs;
    const double *mats;
    ObsNode factor() {
        ObsNode o;
        o.kind = *prog++;
        o.op = *prog++;
        const int nw = *prog++;
        if (nw < 0 || nw > 30) {
            throw InvalidArg("bad observable wire count");
        }
        o.wires.assign(prog, prog + nw);
        prog += nw;
        if (o.kind == 1) {
            if (!mats) {
                throw InvalidArg("Hermitian observable without matrix data");
            }
            const size_t dim = size_t{1} << (2 * nw);
            for (size_t i = 0; i < dim; i++) {
                o.mat.emplace_back(mats[2 * i], mats[2 * i + 1]);
            }
            mats += 2 * dim;
        } else if (o.kind == 0) {
            gate_name(o.op); // range check
        } else {
            throw InvalidArg("bad observable factor kind");
        }
        return o;
    }
    ObsNode term() {
        const int nf = *prog++;
        if (nf < 1) {
            throw InvalidArg("empty observable term");
        }
        if (nf == 1) {
            return factor();
        }
        ObsNode t;
        t.kind = 2;
        std::vector<int> seen;
        for (int f = 0; f < nf; f++) {
            t.kids.push_back(factor());
            for (int w : t.kids.back().wires) {
                if (std::find(seen.begin(), seen.end(), w) != seen.end()) {
                    // TensorProdObsBase ctor, Observables.hpp:238-246
                    throw InvalidArg("All wires in observables must be disjoint.");
                }
                seen.push_back(w);
            }
        }
        return t;
    }
    ObsNode one() {
        const int is_ham = *prog++;
        const int nt = *prog++;
        if (!is_ham) {
            if (nt != 1) {
                throw InvalidArg("non-Hamiltonian observable must have one term");
            }
            return term();
        }
        ObsNode h;
        h.kind = 3;
        for (int t = 0; t < nt; t++) {
            h.coeffs.push_back(*coeffs++);
            h.kids.push_back(term());
        }
        return h;
    }
};

// A Pauli word: (x-mask, zy-mask, #Y) -- the fast path of expval / var /
// the adjoint's observable application.
struct PauliWord {
    uint64_t x = 0, zy = 0;
    int ny = 0;
};

inline bool as_pauli(const ObsNode &o, int n, PauliWord &pw) {
    auto letter = [&](const ObsNode &f) {
        if (f.kind != 0 || f.wires.size() != 1) {
            return false;
        }
        if (f.wires[0] < 0 || f.wires[0] >= n) {
            throw InvalidArg("wire index out of range");
        }
        const uint64_t b = uint64_t{1} << (n - 1 - f.wires[0]);
        switch (f.op) {
        case Identity: return true;
        case PauliX: pw.x |= b; return true;
        case PauliY: pw.x |= b; pw.zy |= b; pw.ny++; return true;
        case PauliZ: pw.zy |= b; return true;
        default: return false;
        }
    };
    pw = PauliWord{};
    if (o.kind == 0) {
        return letter(o);
    }
    if (o.kind == 2) {
        for (const auto &k : o.kids) {
            if (!letter(k)) {
                return false;
            }
        }
        return true;
    }
    return false;
}

template <class F> int guarded(F &&f) {
    try {
        f();
        return PL_OK;
    } catch (const OutOfRange &e) {
        g_err = e.what();
        return PL_ERR_RANGE;
    } catch (const InvalidArg &e) {
        g_err = e.what();
        return PL_ERR_INVALID;
    } catch (const DeviceError &e) {
        g_err = e.what();
        return PL_ERR_DEVICE;
    } catch (const std::exception &e) {
        g_err = e.what();
        return PL_ERR_INVALID;
    }
}

} // namespace plamd

struct pl_sv {
    int n = 0;
    int dtype = PL_C128;
    uint64_t len = 0;
    void *d = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_order = nullptr; // orders work across two handles' streams (pl_sv_copy)
    std::vector<void *> pool;   // scratch states (state-sized)
    size_t pool_used = 0;
    double *red = nullptr;      // reduction scratch (partials)
    size_t red_bytes = 0;
    void *staging = nullptr;    // device staging for matrices / small tables
    size_t staging_bytes = 0;
    std::vector<hipEvent_t> slots; // pl_event_record slots
    std::vector<std::pair<void *, size_t>> bufs; // grow-only device buffers by purpose (plamd::Buf)
    // gates queued while deferral is on (pl_sv_set_deferred), in the tape
    // encoding of pl_sv_apply_ops; applied as tile passes before anything
    // else reads or changes the state
    bool defer = false;
    std::vector<int> q_ids, q_nwires, q_wires, q_nparams;
    std::vector<uint8_t> q_inverse;
    std::vector<double> q_params;

    size_t state_bytes() const { return len * (dtype == PL_C128 ? 16 : 8); }
};

namespace plamd {

inline void *scratch_acquire(pl_sv *h) {
    if (h->pool_used == h->pool.size()) {
        void *p = nullptr;
        HIP_OK(hipMalloc(&p, h->state_bytes()));
        h->pool.push_back(p);
    }
    return h->pool[h->pool_used++];
}
struct ScratchScope {
    pl_sv *h;
    size_t mark;
    explicit ScratchScope(pl_sv *hh) : h(hh), mark(hh->pool_used) {}
    void *get() { return scratch_acquire(h); }
    ~ScratchScope() { h->pool_used = mark; }
};

// Device buffers kept by a handle between calls, one per purpose.
enum Buf { kBufUniforms, kBufIndices, kBufBits, kBufCdfGroups, kBufCdfPrefix, kBufCdfTiles, kBufChainStart,
           kBufScalar, kBufVjp, kBufCount };

// The handle's buffer for `purpose`, at least `bytes` long (grown when
// needed, after the handle's queued work, which may still use the old one).
inline void *buf(pl_sv *h, Buf purpose, size_t bytes) {
    if (h->bufs.size() < kBufCount) {
        h->bufs.resize(kBufCount, {nullptr, 0});
    }
    auto &b = h->bufs[purpose];
    if (b.second < bytes) {
        if (b.first) {
            HIP_OK(hipStreamSynchronize(h->stream));
            HIP_OK(hipFree(b.first));
            b = {nullptr, 0};
        }
        const size_t sz = std::max<size_t>(bytes, 4096);
        HIP_OK(hipMalloc(&b.first, sz));
        b.second = sz;
    }
    return b.first;
}

inline double *reduce_buf(pl_sv *h, size_t bytes) {
    if (h->red_bytes < bytes) {
        if (h->red) {
            HIP_OK(hipStreamSynchronize(h->stream));
            HIP_OK(hipFree(h->red));
            h->red = nullptr;
        }
        HIP_OK(hipMalloc(reinterpret_cast<void **>(&h->red), bytes));
        h->red_bytes = bytes;
    }
    return h->red;
}

// Copy host bytes to the handle's device staging area and wait for the copy,
// so the host buffer may be released as soon as this returns.
inline void *stage(pl_sv *h, const void *src, size_t bytes) {
    if (h->staging_bytes < bytes) {
        HIP_OK(hipStreamSynchronize(h->stream));
        if (h->staging) {
            HIP_OK(hipFree(h->staging));
            h->staging = nullptr;
        }
        const size_t sz = std::max<size_t>(bytes, 1 << 16);
        HIP_OK(hipMalloc(&h->staging, sz));
        h->staging_bytes = sz;
    }
    HIP_OK(hipMemcpyAsync(h->staging, src, bytes, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return h->staging;
}

} // namespace plamd
