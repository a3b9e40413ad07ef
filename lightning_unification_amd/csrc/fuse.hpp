// Gate fusion for the precision path: operations of a tape whose wires
// together span at most `max_wires` wires are multiplied into one dense block
// (2^k x 2^k, k <= 5).  A block is grown from the first operation not yet
// placed by scanning ahead in tape order: an operation joins when its bits
// fit the block and it shares no bit with any operation the scan has passed
// over, so it only moves ahead of operations on disjoint bits, with which it
// commutes (the rule of the tile-pass scheduler).  Each block's matrix is
// built on the host by
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
    std::vector<int> ops;           // tape indices, in application order
    bool fused = true;              // false: apply operation ops[0] unfused
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

constexpr int kFuseLookahead = 512; // operations scanned ahead per block

inline std::vector<FusedBlock> fuse_tape(int n, const TapeView &t, int max_wires, int pad_to) {
    if (max_wires < 1 || max_wires > 5) {
        throw InvalidArg("fusion width must be in [1, 5]");
    }
    pad_to = std::min(pad_to, n);
    std::vector<int> woff(t.nops + 1, 0), poff(t.nops + 1, 0);
    std::vector<uint64_t> mask(t.nops, 0);
    for (int i = 0; i < t.nops; i++) {
        if (t.nwires[i] < 0 || t.nparams[i] < 0) {
            throw InvalidArg("negative wire or parameter count");
        }
        woff[i + 1] = woff[i] + t.nwires[i];
        poff[i + 1] = poff[i] + t.nparams[i];
        if (t.ids[i] == -1) {
            continue;
        }
        for (int a = 0; a < t.nwires[i]; a++) {
            const int w = t.wires[woff[i] + a];
            if (w < 0 || w >= n) {
                throw InvalidArg("wire index out of range");
            }
            mask[i] |= uint64_t{1} << (n - 1 - w);
        }
    }
    std::vector<FusedBlock> out;

    auto build = [&](FusedBlock &b, uint64_t bitmask) {
        for (int p = 0; p < n; p++) {
            if ((bitmask >> p) & 1U) {
                b.bits.push_back(p);
            }
        }
        for (int p = 0; static_cast<int>(b.bits.size()) < pad_to && p < n; p++) {
            if (!((bitmask >> p) & 1U)) {
                b.bits.push_back(p);
            }
        }
        std::sort(b.bits.begin(), b.bits.end());
        const int k = static_cast<int>(b.bits.size());
        const size_t dim = size_t{1} << k;
        std::vector<GroupOp<double>> ops;
        for (int i : b.ops) {
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
    };

    std::vector<char> placed(t.nops, 0);
    int pos = 0;
    while (true) {
        while (pos < t.nops && (placed[pos] || t.ids[pos] == -1)) { // state preparation: skipped
            pos++;
        }
        if (pos >= t.nops) {
            break;
        }
        placed[pos] = 1;
        FusedBlock b;
        b.ops.push_back(pos);
        if (t.nwires[pos] > max_wires) { // applied unfused
            b.fused = false;
            out.push_back(std::move(b));
            continue;
        }
        uint64_t cur = mask[pos], blocked = 0;
        const uint64_t all = n >= 64 ? ~uint64_t{0} : (uint64_t{1} << n) - 1;
        for (int j = pos + 1; j < t.nops && j <= pos + kFuseLookahead && blocked != all; j++) {
            if (placed[j] || t.ids[j] == -1) {
                continue;
            }
            const uint64_t u = cur | mask[j];
            if (t.nwires[j] <= max_wires && !(mask[j] & blocked) && __builtin_popcountll(u) <= max_wires) {
                cur = u;
                placed[j] = 1;
                b.ops.push_back(j);
            } else {
                blocked |= mask[j];
            }
        }
        build(b, cur);
        out.push_back(std::move(b));
    }
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
