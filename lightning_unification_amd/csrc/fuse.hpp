// Gate fusion for the precision path: consecutive operations of a tape whose
// wires together span at most `max_wires` wires are multiplied into one dense
// block (2^k x 2^k, k <= 5).  Each block's matrix is built on the host by
// applying the very group ops the GPU kernels apply (lower_gate) to the 2^k
// basis vectors of a k-qubit local system, so fused and unfused runs share
// one definition of every gate.  An operation on more than `max_wires` wires
// stays a block of its own and is applied unfused.
//
// Block-local index j: bit m of j is the amplitude's bit at position bits[m]
// (bits ascending).  Blocks narrower than `pad_to` are widened with the lowest
// bit positions they do not touch, acting as the identity there, so every
// fused launch can use one 5-bit kernel.
#pragma once

#include <algorithm>
#include <complex>
#include <vector>

#include "lower.hpp"

namespace plamd {

struct FusedBlock {
    int first_op = 0;
    int nops = 0;
    bool fused = true;              // false: apply operation first_op unfused
    std::vector<int> bits;          // ascending bit positions
    std::vector<std::complex<double>> U; // 2^k x 2^k row-major in the block-local index
};

// v <- g v on a k-qubit vector (the semantics of the gate kernels)
template <class P> void host_apply_group(const GroupOp<P> &g, std::complex<P> *v, int k) {
    const uint64_t len = uint64_t{1} << k;
    if (g.kind == GroupOp<P>::Noop) {
        return;
    }
    if (g.kind == GroupOp<P>::Parity) {
        for (uint64_t i = 0; i < len; i++) {
            v[i] *= g.parity_phase[__builtin_popcountll(i & g.parity_mask) & 1];
        }
        return;
    }
    const int M = g.M;
    std::vector<std::complex<P>> x(M);
    for (uint64_t q = 0; q < (len >> g.nb); q++) {
        uint64_t base = q;
        for (int b = 0; b < g.nb; b++) {
            const int p = g.pos[b];
            base = ((base >> p) << (p + 1)) | (base & ((uint64_t{1} << p) - 1));
        }
        base |= g.fixed;
        for (int j = 0; j < M; j++) {
            x[j] = v[base | g.off[j]];
        }
        for (int r = 0; r < M; r++) {
            std::complex<P> acc = 0;
            for (int c = 0; c < M; c++) {
                acc += g.mat[r * M + c] * x[c];
            }
            v[base | g.off[r]] = acc;
        }
    }
}

struct TapeView {
    int nops;
    const int *ids, *nwires, *wires;
    const uint8_t *inverse;
    const int *nparams;
    const double *params;
};

inline std::vector<FusedBlock> fuse_tape(int n, const TapeView &t, int max_wires, int pad_to) {
    if (max_wires < 1 || max_wires > 5) {
        throw InvalidArg("fusion width must be in [1, 5]");
    }
    pad_to = std::min(pad_to, n);
    std::vector<int> woff(t.nops + 1, 0), poff(t.nops + 1, 0);
    for (int i = 0; i < t.nops; i++) {
        if (t.nwires[i] < 0 || t.nparams[i] < 0) {
            throw InvalidArg("negative wire or parameter count");
        }
        woff[i + 1] = woff[i] + t.nwires[i];
        poff[i + 1] = poff[i] + t.nparams[i];
    }
    std::vector<FusedBlock> out;
    std::vector<int> cur; // bit positions of the open block
    int start = 0;

    auto close = [&](int end) {
        if (end <= start) {
            return;
        }
        FusedBlock b;
        b.first_op = start;
        b.nops = end - start;
        b.bits = cur;
        for (int p = 0; static_cast<int>(b.bits.size()) < pad_to && p < n; p++) {
            if (std::find(b.bits.begin(), b.bits.end(), p) == b.bits.end()) {
                b.bits.push_back(p);
            }
        }
        std::sort(b.bits.begin(), b.bits.end());
        const int k = static_cast<int>(b.bits.size());
        const size_t dim = size_t{1} << k;
        std::vector<GroupOp<double>> ops;
        for (int i = start; i < end; i++) {
            if (t.ids[i] == -1) {
                continue; // state preparation: skipped, as by applyOperations
            }
            std::vector<int> lw(t.nwires[i]);
            for (int a = 0; a < t.nwires[i]; a++) {
                const int bit = n - 1 - t.wires[woff[i] + a];
                const int m = static_cast<int>(std::find(b.bits.begin(), b.bits.end(), bit) - b.bits.begin());
                lw[a] = k - 1 - m; // local wire of local bit m
            }
            ops.push_back(lower_gate<double>(k, t.ids[i], lw.data(), t.nwires[i], t.inverse[i] != 0,
                                             t.params + poff[i], t.nparams[i]));
        }
        b.U.assign(dim * dim, 0);
        std::vector<std::complex<double>> col(dim);
        for (size_t c = 0; c < dim; c++) {
            std::fill(col.begin(), col.end(), 0);
            col[c] = 1;
            for (const auto &g : ops) {
                host_apply_group(g, col.data(), k);
            }
            for (size_t r = 0; r < dim; r++) {
                b.U[r * dim + c] = col[r];
            }
        }
        out.push_back(std::move(b));
        cur.clear();
    };

    for (int i = 0; i < t.nops; i++) {
        if (t.ids[i] == -1) {
            continue;
        }
        std::vector<int> bits;
        for (int a = 0; a < t.nwires[i]; a++) {
            const int w = t.wires[woff[i] + a];
            if (w < 0 || w >= n) {
                throw InvalidArg("wire index out of range");
            }
            bits.push_back(n - 1 - w);
        }
        if (static_cast<int>(bits.size()) > max_wires) {
            close(i);
            FusedBlock b;
            b.first_op = i;
            b.nops = 1;
            b.fused = false;
            out.push_back(std::move(b));
            start = i + 1;
            continue;
        }
        std::vector<int> uni = cur;
        for (int b : bits) {
            if (std::find(uni.begin(), uni.end(), b) == uni.end()) {
                uni.push_back(b);
            }
        }
        if (static_cast<int>(uni.size()) > max_wires) {
            close(i);
            start = i;
            uni = bits;
            std::sort(uni.begin(), uni.end());
            uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
        }
        cur = uni;
    }
    close(t.nops);
    return out;
}

// Member offsets of a block: local index j -> OR of 1 << bits[m] over set bits m of j.
inline std::vector<uint64_t> block_offsets(const std::vector<int> &bits) {
    std::vector<uint64_t> off(size_t{1} << bits.size(), 0);
    for (size_t j = 0; j < off.size(); j++) {
        for (size_t m = 0; m < bits.size(); m++) {
            if ((j >> m) & 1U) {
                off[j] |= uint64_t{1} << bits[m];
            }
        }
    }
    return off;
}

} // namespace plamd
