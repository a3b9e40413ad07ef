// Gate fusion for the precision path: operations of a tape whose wires
// together span at most `max_wires` wires are multiplied into one dense block
// (2^k x 2^k, k <= 5).  A block has a wire set; scanning ahead in tape order
// from the first operation not yet placed, an operation joins when its bits
// lie in the set and it shares no bit with any operation the scan has passed
// over, so it only moves ahead of operations on disjoint bits, with which it
// commutes (the rule of the tile-pass scheduler).  The set starts as the
// wires claimed in tape order and is improved by exchanging wires while the
// block takes more operations.  Each block's matrix is
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

constexpr int kFuseLookahead = 512;  // operations scanned ahead per block
constexpr int kFuseSearchRounds = 16; // wire-set exchanges tried per block, at most
constexpr int kFuseDeepQubits = 28;   // states from which two-wire exchanges are tried too

// The blocks of a tape, one at a time (the executor builds and launches a
// block while it plans the next).
class FuseScheduler {
  public:
    FuseScheduler(int n, const TapeView &t, int max_wires, int pad_to)
        : n_(n), t_(t), max_wires_(max_wires), pad_to_(std::min(pad_to, n)), woff_(t.nops + 1, 0),
          poff_(t.nops + 1, 0), mask_(t.nops, 0), placed_(t.nops, 0) {
        if (max_wires < 1 || max_wires > 5) {
            throw InvalidArg("fusion width must be in [1, 5]");
        }
        for (int i = 0; i < t.nops; i++) {
            if (t.nwires[i] < 0 || t.nparams[i] < 0) {
                throw InvalidArg("negative wire or parameter count");
            }
            woff_[i + 1] = woff_[i] + t.nwires[i];
            poff_[i + 1] = poff_[i] + t.nparams[i];
            if (t.ids[i] == -1) {
                continue;
            }
            for (int a = 0; a < t.nwires[i]; a++) {
                const int w = t.wires[woff_[i] + a];
                if (w < 0 || w >= n) {
                    throw InvalidArg("wire index out of range");
                }
                mask_[i] |= uint64_t{1} << (n - 1 - w);
            }
        }
    }

    // The next block, its matrix built (false once every operation is placed).
    bool next(FusedBlock &b) {
        const TapeView &t = t_;
        const int n = n_, max_wires = max_wires_;
        while (pos_ < t.nops && (placed_[pos_] || t.ids[pos_] == -1)) { // state preparation: skipped
            pos_++;
        }
        if (pos_ >= t.nops) {
            return false;
        }
        const int pos = pos_;
        b = FusedBlock{};
        if (t.nwires[pos] > max_wires) { // applied unfused
            placed_[pos] = 1;
            b.ops.push_back(pos);
            b.fused = false;
            return true;
        }
        const int end = std::min(t.nops, pos + kFuseLookahead + 1);
        uint64_t all = 0; // every bit the scanned operations touch
        for (int j = pos; j < end; j++) {
            if (!placed_[j] && t.ids[j] != -1) {
                all |= mask_[j];
            }
        }
        // the operations a block on wire set `set` takes, in order: each one
        // inside the set that shares no bit with an operation passed over
        auto absorb = [&](uint64_t set, std::vector<int> *take) {
            uint64_t blocked = 0;
            int taken = 0;
            for (int j = pos; j < end && (blocked & all) != all; j++) {
                if (placed_[j] || t.ids[j] == -1) {
                    continue;
                }
                if (t.nwires[j] <= max_wires && !(mask_[j] & blocked) && (mask_[j] & ~set) == 0) {
                    taken++;
                    if (take) {
                        take->push_back(j);
                    }
                } else {
                    blocked |= mask_[j];
                }
            }
            return taken;
        };
        // first guess: wires claimed in tape order from the first unplaced
        // operation on
        uint64_t set = mask_[pos];
        {
            uint64_t blocked = 0;
            for (int j = pos + 1; j < end && (blocked & all) != all; j++) {
                if (placed_[j] || t.ids[j] == -1) {
                    continue;
                }
                const uint64_t u = set | mask_[j];
                if (t.nwires[j] <= max_wires && !(mask_[j] & blocked) && __builtin_popcountll(u) <= max_wires) {
                    set = u;
                } else {
                    blocked |= mask_[j];
                }
            }
        }
        // then add or exchange wires while the block takes more operations
        // (the rule of the tile-pass scheduler; the first operation need not
        // stay: one left out opens a later block).  On the 50-layer random
        // circuit of 33 qubits: 185 blocks, 163 with single exchanges, 148
        // with pairs.
        int best = absorb(set, nullptr);
        for (int round = 0; round < kFuseSearchRounds; round++) {
            uint64_t best_set = set;
            int best_n = best;
            for (int b2 = 0; b2 < n; b2++) {
                if ((set >> b2) & 1U) {
                    continue;
                }
                const uint64_t in = uint64_t{1} << b2;
                if (__builtin_popcountll(set) < max_wires) {
                    const int c = absorb(set | in, nullptr);
                    if (c > best_n) {
                        best_n = c;
                        best_set = set | in;
                    }
                }
                for (int a = 0; a < n; a++) {
                    if ((set >> a) & 1U) {
                        const uint64_t m = (set & ~(uint64_t{1} << a)) | in;
                        const int c = absorb(m, nullptr);
                        if (c > best_n) {
                            best_n = c;
                            best_set = m;
                        }
                    }
                }
            }
            if (best_n == best && n >= kFuseDeepQubits) { // the first two-wire exchange that helps
                for (int a = 0; a < n && best_n == best; a++) {
                    for (int a2 = a + 1; a2 < n && best_n == best && ((set >> a) & 1U); a2++) {
                        if (!((set >> a2) & 1U)) {
                            continue;
                        }
                        for (int c1 = 0; c1 < n && best_n == best; c1++) {
                            for (int c2 = c1 + 1; c2 < n && !((set >> c1) & 1U); c2++) {
                                if ((set >> c2) & 1U) {
                                    continue;
                                }
                                const uint64_t m = (set & ~(uint64_t{1} << a) & ~(uint64_t{1} << a2)) |
                                                   (uint64_t{1} << c1) | (uint64_t{1} << c2);
                                const int c = absorb(m, nullptr);
                                if (c > best) {
                                    best_n = c;
                                    best_set = m;
                                    break;
                                }
                            }
                        }
                    }
                }
            }
            if (best_n == best) {
                break;
            }
            best = best_n;
            set = best_set;
        }
        absorb(set, &b.ops);
        uint64_t used = 0; // the block's width: the wires its operations touch
        for (int j : b.ops) {
            placed_[j] = 1;
            used |= mask_[j];
        }
        build(b, used);
        return true;
    }

  private:
    void build(FusedBlock &b, uint64_t bitmask) const {
        const TapeView &t = t_;
        const int n = n_;
        for (int p = 0; p < n; p++) {
            if ((bitmask >> p) & 1U) {
                b.bits.push_back(p);
            }
        }
        for (int p = 0; static_cast<int>(b.bits.size()) < pad_to_ && p < n; p++) {
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
                const int bit = n - 1 - t.wires[woff_[i] + a];
                const int m = static_cast<int>(std::find(b.bits.begin(), b.bits.end(), bit) - b.bits.begin());
                lw[a] = k - 1 - m; // local wire of local bit m
            }
            ops.push_back(lower_gate<double>(k, t.ids[i], lw.data(), t.nwires[i], t.inverse[i] != 0,
                                             t.params + poff_[i], t.nparams[i]));
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
    }

    int n_;
    TapeView t_;
    int max_wires_, pad_to_;
    std::vector<int> woff_, poff_;
    std::vector<uint64_t> mask_;
    std::vector<char> placed_;
    int pos_ = 0;
};

inline std::vector<FusedBlock> fuse_tape(int n, const TapeView &t, int max_wires, int pad_to) {
    FuseScheduler s(n, t, max_wires, pad_to);
    std::vector<FusedBlock> out;
    FusedBlock b;
    while (s.next(b)) {
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
