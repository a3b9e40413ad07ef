// Tile-pass scheduling (host only): groups the operations of a tape into
// passes of k_tile_pass.
//
// A pass owns the `low` lowest bit positions and k - low higher ones.  Given
// those high bits, operations are taken in tape order; an operation joins the
// pass when
//   - it lowers to a group op of at most kTpMaxM amplitudes and kTpMaxPins
//     pins (or to a parity phase),
//   - its high bits are among the pass's, and
//   - it shares no bit with any earlier operation left out of the pass.
// The last rule makes the reordering exact: an operation only moves ahead of
// operations on disjoint bits, with which it commutes.  Operations left out
// keep their relative order and are scheduled into later passes.  An
// operation that can never be tiled (too many amplitudes per group) becomes a
// pass of its own, applied by the per-gate kernels.  The high bits start as
// those the operations claim in tape order (padded with the lowest unused
// positions) and are then improved by exchanging bits while the pass takes
// more operations.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "fuse.hpp"
#include "kernels_tile.hpp"
#include "lower.hpp"

namespace plamd {

constexpr size_t kTileLookahead = 4096; // operations scanned per pass
constexpr int kTileSearchRounds = 16;    // bit exchanges tried per pass, at most
constexpr int kTileDeepSearchQubits = 30; // states from which pair exchanges are tried too

struct TilePass {
    bool tiled = true;        // false: ops[0] applied by the per-gate kernels
    std::vector<int> hb;      // bit positions of local bits low.. (ascending; tile_lane_order reorders them)
    std::vector<int> ops;     // tape op indices in application order
};

// Does the operation lower to something a tile pass can apply?
inline bool tile_op_ok(int id, int nw, bool inv, const double *prm, int np) {
    if (nw < 1 || nw > kTpMaxPins) {
        return false;
    }
    std::vector<int> lw(nw);
    for (int a = 0; a < nw; a++) {
        lw[a] = a;
    }
    const auto g = lower_gate<double>(nw, id, lw.data(), nw, inv, prm, np);
    if (g.kind == GroupOp<double>::Noop || g.kind == GroupOp<double>::Parity) {
        return true;
    }
    return (g.M == 1 || g.M == 2 || g.M == 4 || g.M == 8) && g.nb <= kTpMaxPins;
}

inline void tile_check_args(int n, int k, int low) {
    if (k < kTpMinBits || k > kTpMaxBits) {
        throw InvalidArg("tile bits must be in [9, 12]");
    }
    if (low < 0 || low > kTpMaxLow) {
        throw InvalidArg("low bits must be in [0, 6]");
    }
    if (n < k) {
        throw InvalidArg("tile bits exceed the qubit count");
    }
}

// The passes of a tape, one at a time (the executor plans a pass while the
// device applies the one before).
class TileScheduler {
  public:
    TileScheduler(int n, const TapeView &t, int k, int low)
        : n_((tile_check_args(n, k, low), n)), low_(low), budget_(k - low), low_mask_((uint64_t{1} << low) - 1),
          all_(0), deep_(n >= kTileDeepSearchQubits),
          bits_(t.nops, 0), ok_(t.nops, 0) {
        std::vector<int> woff(t.nops + 1, 0), poff(t.nops + 1, 0);
        for (int i = 0; i < t.nops; i++) {
            if (t.nwires[i] < 0 || t.nparams[i] < 0) {
                throw InvalidArg("negative wire or parameter count");
            }
            woff[i + 1] = woff[i] + t.nwires[i];
            poff[i + 1] = poff[i] + t.nparams[i];
        }
        for (int i = 0; i < t.nops; i++) {
            if (t.ids[i] == -1) {
                continue; // state preparation: skipped, as by applyOperations
            }
            for (int a = 0; a < t.nwires[i]; a++) {
                const int w = t.wires[woff[i] + a];
                if (w < 0 || w >= n) {
                    throw InvalidArg("wire index out of range");
                }
                bits_[i] |= uint64_t{1} << (n - 1 - w);
            }
            ok_[i] = tile_op_ok(t.ids[i], t.nwires[i], t.inverse[i] != 0, t.params + poff[i], t.nparams[i]) &&
                     __builtin_popcountll(bits_[i] & ~low_mask_) <= budget_;
            pending_.push_back(i);
        }
    }

    // The next pass (false once every operation is scheduled).
    bool next(TilePass &p) {
        if (pending_.empty()) {
            return false;
        }
        p = TilePass{};
        // every bit the scanned operations touch: once all of them are
        // blocked, no later operation can join
        all_ = 0;
        for (size_t j = 0; j < pending_.size() && j < kTileLookahead; j++) {
            all_ |= bits_[pending_[j]];
        }
        // first guess: high bits claimed in tape order until the budget is spent
        uint64_t hmask = 0, blocked = 0;
        for (size_t j = 0; j < pending_.size() && j < kTileLookahead && blocked != all_; j++) {
            const int i = pending_[j];
            const uint64_t hi = bits_[i] & ~low_mask_;
            if (ok_[i] && !(bits_[i] & blocked) && __builtin_popcountll(hmask | hi) <= budget_) {
                hmask |= hi;
            } else {
                blocked |= bits_[i];
            }
        }
        for (int b = low_; b < n_ && __builtin_popcountll(hmask) < budget_; b++) {
            hmask |= uint64_t{1} << b;
        }
        // then exchange high bits for others while that lets the pass take
        // more operations (a bit claimed early can cost the pass many later
        // ones): the best single exchange each round, or failing one (on
        // states of kTileDeepSearchQubits or more, where a pass takes the
        // device longer than this search takes the host) the first pair
        // exchange that helps.  On a 50-layer random circuit of 30 qubits
        // (2^11-amplitude tiles) the passes go from 81 to 62 with single
        // exchanges and to 55 with pairs.
        int best = absorb(hmask, nullptr, nullptr);
        for (int round = 0; round < kTileSearchRounds && best > 0; round++) {
            uint64_t best_mask = hmask;
            int best_n = best;
            for (int a = low_; a < n_; a++) {
                if (!((hmask >> a) & 1U)) {
                    continue;
                }
                for (int b = low_; b < n_; b++) {
                    if ((hmask >> b) & 1U) {
                        continue;
                    }
                    const uint64_t m = (hmask & ~(uint64_t{1} << a)) | (uint64_t{1} << b);
                    const int c = absorb(m, nullptr, nullptr);
                    if (c > best_n) {
                        best_n = c;
                        best_mask = m;
                    }
                }
            }
            if (best_n == best && deep_) {
                best_mask = pair_exchange(hmask, best, best_n);
            }
            if (best_n == best) {
                break;
            }
            best = best_n;
            hmask = best_mask;
        }
        std::vector<int> rest;
        absorb(hmask, &p.ops, &rest);
        if (p.ops.empty()) { // the first pending op cannot be tiled
            p.tiled = false;
            p.ops.push_back(pending_[0]);
            rest.erase(rest.begin());
        } else {
            for (int b = low_; b < n_; b++) {
                if ((hmask >> b) & 1U) {
                    p.hb.push_back(b);
                }
            }
        }
        pending_.swap(rest);
        return true;
    }

  private:
    // The pending operations a pass with high bits `hmask` takes, in order:
    // each op whose high bits lie in hmask and which shares no bit with an op
    // left out before it.  Fills `take` and `left` when given.
    int absorb(uint64_t hmask, std::vector<int> *take, std::vector<int> *left) const {
        uint64_t blocked = 0;
        int taken = 0;
        size_t j = 0;
        for (; j < pending_.size() && j < kTileLookahead; j++) {
            const int i = pending_[j];
            if (ok_[i] && !(bits_[i] & blocked) && ((bits_[i] & ~low_mask_) & ~hmask) == 0) {
                taken++;
                if (take) {
                    take->push_back(i);
                }
            } else {
                blocked |= bits_[i];
                if (left) {
                    left->push_back(i);
                }
                if ((blocked & all_) == all_) {
                    j++;
                    break;
                }
            }
        }
        if (left) {
            left->insert(left->end(), pending_.begin() + static_cast<long>(j), pending_.end());
        }
        return taken;
    }

    // The first exchange of two high bits for two others that takes more
    // than `best` operations (its count in best_n), else hmask unchanged.
    uint64_t pair_exchange(uint64_t hmask, int best, int &best_n) const {
        for (int a = low_; a < n_; a++) {
            for (int a2 = a + 1; a2 < n_ && ((hmask >> a) & 1U); a2++) {
                if (!((hmask >> a2) & 1U)) {
                    continue;
                }
                for (int b = low_; b < n_; b++) {
                    for (int b2 = b + 1; b2 < n_ && !((hmask >> b) & 1U); b2++) {
                        if ((hmask >> b2) & 1U) {
                            continue;
                        }
                        const uint64_t m = (hmask & ~(uint64_t{1} << a) & ~(uint64_t{1} << a2)) |
                                           (uint64_t{1} << b) | (uint64_t{1} << b2);
                        const int c = absorb(m, nullptr, nullptr);
                        if (c > best) {
                            best_n = c;
                            return m;
                        }
                    }
                }
            }
        }
        return hmask;
    }

    int n_, low_, budget_;
    uint64_t low_mask_, all_; // all_: the bits of the operations the current pass may take
    bool deep_;
    std::vector<uint64_t> bits_;
    std::vector<char> ok_;
    std::vector<int> pending_;
};

inline std::vector<TilePass> tile_schedule(int n, const TapeView &t, int k, int low) {
    TileScheduler s(n, t, k, low);
    std::vector<TilePass> out;
    TilePass p;
    while (s.next(p)) {
        out.push_back(std::move(p));
    }
    return out;
}

// Operation i of the tape lowered onto the k-qubit local system of a pass.
template <class T>
GroupOp<T> tile_group(int n, const TapeView &t, int i, int woff, int poff, int low, const std::vector<int> &hb) {
    const int k = low + static_cast<int>(hb.size());
    std::vector<int> lw(t.nwires[i]);
    for (int a = 0; a < t.nwires[i]; a++) {
        const int bit = n - 1 - t.wires[woff + a];
        int m = bit;
        if (bit >= low) {
            const auto it = std::find(hb.begin(), hb.end(), bit);
            if (it == hb.end()) {
                throw InvalidArg("operation outside its tile");
            }
            m = low + static_cast<int>(it - hb.begin());
        }
        lw[a] = k - 1 - m;
    }
    std::vector<T> prm(t.nparams[i]);
    for (int a = 0; a < t.nparams[i]; a++) {
        prm[a] = static_cast<T>(t.params[poff + a]);
    }
    return lower_gate<T>(k, t.ids[i], lw.data(), t.nwires[i], t.inverse[i] != 0, prm.data(), t.nparams[i]);
}

template <class T> TpOp<T> tile_lds_op(const GroupOp<T> &g) {
    TpOp<T> op{};
    if (g.kind == GroupOp<T>::Parity) {
        op.kind = 2;
        op.M = 1;
        op.mask = static_cast<uint32_t>(g.parity_mask);
        op.mat[0] = {g.parity_phase[0].real(), g.parity_phase[0].imag()};
        op.mat[1] = {g.parity_phase[1].real(), g.parity_phase[1].imag()};
        return op;
    }
    if (!(g.M == 1 || g.M == 2 || g.M == 4 || g.M == 8) || g.nb > kTpMaxPins) {
        throw InvalidArg("operation too wide for a tile pass");
    }
    op.kind = g.diag ? 1 : 0;
    op.M = g.M;
    op.nb = g.nb;
    for (int a = 0; a < g.nb; a++) {
        op.pos[a] = g.pos[a];
    }
    op.fixed = static_cast<uint32_t>(g.fixed);
    for (int a = 0; a < g.M; a++) {
        op.off[a] = static_cast<uint32_t>(g.off[a]);
    }
    for (int a = 0; a < g.M * g.M; a++) {
        op.mat[a] = {g.mat[a].real(), g.mat[a].imag()};
    }
    return op;
}

// Compiles one tiled pass into a wavefront-tile program (kernels_tile.hpp):
// lane bits are local bits 0..5, register bit j is local bit 6 + j.
// A single-bit operation as (bit, 2x2 matrix): a 2-member group on one bit
// with nothing else pinned, or a phase on one bit (a 1-member group whose
// member is the bit set).  Returns -1 for anything else.
template <class T> int single_bit(const GroupOp<T> &g, std::complex<T> (&m)[4]) {
    if (g.kind != GroupOp<T>::Group || g.nb != 1) {
        return -1;
    }
    const int b = g.pos[0];
    const uint64_t bm = uint64_t{1} << b;
    if (g.M == 2 && g.fixed == 0 && g.off[0] == 0 && g.off[1] == bm) {
        for (int j = 0; j < 4; j++) {
            m[j] = g.mat[j];
        }
        return b;
    }
    if (g.M == 1 && g.fixed == bm && g.off[0] == 0) {
        m[0] = 1;
        m[1] = m[2] = 0;
        m[3] = g.mat[0];
        return b;
    }
    return -1;
}

// Pinned and member bits of a group operation, and whether its fixed bits
// are all pinned non-member bits (the forms the tile program applies).
template <class T> bool tile_clean(const GroupOp<T> &g, uint32_t &pins, uint32_t &xo) {
    pins = xo = 0;
    for (int b = 0; b < g.nb; b++) {
        pins |= 1U << g.pos[b];
    }
    for (int b = 0; b < g.M; b++) {
        xo |= static_cast<uint32_t>(g.off[b]);
    }
    const uint32_t fx = static_cast<uint32_t>(g.fixed);
    return (fx & ~pins) == 0 && (fx & xo) == 0 && (xo & ~pins) == 0;
}

// A diagonal group operation pinning at most two bits (its phases fold into
// one table over those bits).
template <class T> bool tile_small_diag(const GroupOp<T> &g) {
    uint32_t pins, xo;
    return g.kind == GroupOp<T>::Group && (g.M == 1 || g.diag) && g.M <= 4 && tile_clean(g, pins, xo) &&
           __builtin_popcount(pins) <= 2;
}

// The phase such an operation applies to local index i (1 outside its groups).
template <class T> std::complex<T> tile_diag_phase(const GroupOp<T> &g, uint32_t i) {
    uint32_t pins, xo;
    tile_clean(g, pins, xo);
    const uint32_t cmask = pins & ~xo;
    if ((i & cmask) != (static_cast<uint32_t>(g.fixed) & cmask)) {
        return 1;
    }
    for (int j = 0; j < g.M; j++) {
        if (static_cast<uint32_t>(g.off[j]) == (i & xo)) {
            return g.mat[j * g.M + j];
        }
    }
    return 1;
}

// The pass's operations lowered to the tile's local bits, in application
// order, with two kinds of merges:
//  - single-bit operations on the same bit: an operation on bit b whose last
//    predecessor touching b is also a single-bit operation on b is multiplied
//    into it (the operations between touch other bits and commute with both);
//  - diagonal operations: one on at most two bits is multiplied into the
//    last earlier diagonal one it can move back to (every operation between
//    them that touches its bits is diagonal too, so all of them commute)
//    when the two pin at most two bits together.
// A random circuit's pass loses about a fifth of its operations to the first
// rule and a sixth of its diagonal ones to the second.
template <class T>
std::vector<GroupOp<T>> tile_ops_merged(int n, const TapeView &t, const TilePass &ps, int low,
                                        const std::vector<int> &woff, const std::vector<int> &poff) {
    const int k = low + static_cast<int>(ps.hb.size());
    std::vector<GroupOp<T>> out;
    std::vector<int> last(k, -1); // index in `out` of the last op touching each local bit
    auto touch_of = [](const GroupOp<T> &g) {
        uint64_t touch = 0;
        if (g.kind == GroupOp<T>::Parity) {
            touch = g.parity_mask;
        } else {
            for (int j = 0; j < g.nb; j++) {
                touch |= uint64_t{1} << g.pos[j];
            }
        }
        return touch;
    };
    auto diagonal = [](const GroupOp<T> &g) {
        return g.kind == GroupOp<T>::Parity || (g.kind == GroupOp<T>::Group && (g.M == 1 || g.diag));
    };
    for (int i : ps.ops) {
        auto g = tile_group<T>(n, t, i, woff[i], poff[i], low, ps.hb);
        if (g.kind == GroupOp<T>::Noop) {
            continue;
        }
        std::complex<T> m[4], p[4];
        const int b = single_bit(g, m);
        if (b >= 0 && last[b] >= 0 && single_bit(out[last[b]], p) == b) {
            // out[last[b]] <- g * out[last[b]]
            std::vector<std::complex<T>> mm = {m[0] * p[0] + m[1] * p[2], m[0] * p[1] + m[1] * p[3],
                                               m[2] * p[0] + m[3] * p[2], m[2] * p[1] + m[3] * p[3]};
            const bool dg = mm[1] == std::complex<T>(0) && mm[2] == std::complex<T>(0);
            out[last[b]] = make_group<T>({b}, 0, {0, uint64_t{1} << b}, std::move(mm), dg);
            continue;
        }
        const uint64_t touch = touch_of(g);
        if (tile_small_diag(g)) {
            int into = -1;
            for (int j = static_cast<int>(out.size()) - 1; j >= 0; j--) {
                const uint64_t tj = touch_of(out[j]);
                if (tile_small_diag(out[j]) && __builtin_popcountll(tj | touch) <= 2) {
                    into = j;
                    break;
                }
                if ((tj & touch) && !diagonal(out[j])) {
                    break;
                }
            }
            if (into >= 0) {
                // one table over the union of the two operations' bits
                const uint64_t u = touch_of(out[into]) | touch;
                std::vector<int> bits;
                for (int j = 0; j < k; j++) {
                    if ((u >> j) & 1U) {
                        bits.push_back(j);
                    }
                }
                const int M = 1 << bits.size();
                std::vector<uint64_t> off(M, 0);
                std::vector<std::complex<T>> mm(static_cast<size_t>(M) * M, std::complex<T>(0));
                for (int c = 0; c < M; c++) {
                    for (size_t a = 0; a < bits.size(); a++) {
                        off[c] |= ((c >> a) & 1) ? uint64_t{1} << bits[a] : 0;
                    }
                    const uint32_t li = static_cast<uint32_t>(off[c]);
                    mm[c * M + c] = tile_diag_phase(g, li) * tile_diag_phase(out[into], li);
                }
                out[into] = make_group<T>(bits, 0, std::move(off), std::move(mm), true);
                for (int j : bits) {
                    last[j] = std::max(last[j], into);
                }
                continue;
            }
        }
        out.push_back(std::move(g));
        for (int j = 0; j < k; j++) {
            if ((touch >> j) & 1U) {
                last[j] = static_cast<int>(out.size()) - 1;
            }
        }
    }
    return out;
}

// Relative cost of an operation in a tile program when the local bits in
// `lanes` are the lane bits, following the choices of tile_program.  Weights:
// the per-operation times fitted to a pass over 2^30 complex128 amplitudes
// in 2^11-amplitude tiles (ms): diagonal 0.15, 2x2 on a register bit 0.12,
// register pair 0.11, 4x4 on two register bits 0.54, 2x2 on a lane bit 0.15,
// lane x register 4x4 0.43, 4x4 on two lane bits 0.55; LDS about 1.
template <class T> double tile_op_cost(const GroupOp<T> &g, uint32_t lanes) {
    if (g.kind == GroupOp<T>::Parity) {
        return 0.15;
    }
    uint32_t pins, xo;
    const bool clean = tile_clean(g, pins, xo);
    const int nx = __builtin_popcount(xo);
    if (clean && (g.M == 1 || g.diag) && g.M <= 4 && nx <= 2) {
        return 0.15;
    }
    const bool p1 = clean && g.M == 2 && g.off[0] == 0 && nx == 1;
    const bool p2 = clean && g.M == 2 && nx == 2 && __builtin_popcountll(g.off[0]) == 1 &&
                    __builtin_popcountll(g.off[1]) == 1;
    const bool p4 = clean && g.M == 4 && g.off[0] == 0 && nx == 2 && __builtin_popcountll(g.off[1]) == 1 &&
                    __builtin_popcountll(g.off[2]) == 1 && g.off[3] == (g.off[1] | g.off[2]);
    const bool lane_cond = (pins & ~xo & lanes) != 0;
    if ((p1 || p2 || (p4 && !lane_cond)) && (xo & lanes) == 0) {
        return p1 ? 0.12 : (p2 ? 0.11 : 0.54);
    }
    if (p1) {
        return 0.15;
    }
    if (clean && g.kind == GroupOp<T>::Group && g.M <= 4 && nx == 2 && (xo & lanes) != 0) {
        return __builtin_popcount(xo & lanes) == 1 ? 0.43 : 0.55;
    }
    return 1.0;
}

// The order of a tiled pass's high bits that gives the cheapest program: the
// 6 - low lane bits above the low ones are the high bits that minimise the
// summed tile_op_cost (ties: the lowest such bits), then the register bits,
// each part ascending.  Every order is exact (the local bits only place the
// amplitudes in the wavefront tile); the search is skipped where it would
// weigh more than 512 choices.
template <class T>
std::vector<int> tile_lane_order(int n, const TapeView &t, const TilePass &ps, int low, const std::vector<int> &woff,
                                 const std::vector<int> &poff) {
    const int nh = static_cast<int>(ps.hb.size()), nfree = std::max(0, 6 - low);
    if (!ps.tiled || nfree == 0 || nfree >= nh) {
        return ps.hb;
    }
    double choices = 1;
    for (int i = 0; i < nfree; i++) {
        choices = choices * (nh - i) / (i + 1);
    }
    if (choices > 512) {
        return ps.hb;
    }
    const auto ops = tile_ops_merged<T>(n, t, ps, low, woff, poff);
    const uint32_t lowlanes = (1U << low) - 1;
    uint32_t best_sel = (1U << nfree) - 1;
    double best = 0;
    for (const auto &g : ops) {
        best += tile_op_cost(g, lowlanes | (best_sel << low));
    }
    for (uint32_t sel = 0; sel < (1U << nh); sel++) {
        if (__builtin_popcount(sel) != nfree || sel == ((1U << nfree) - 1)) {
            continue;
        }
        double c = 0;
        for (const auto &g : ops) {
            c += tile_op_cost(g, lowlanes | (sel << low));
        }
        if (c < best - 1e-9) {
            best = c;
            best_sel = sel;
        }
    }
    std::vector<int> hb;
    for (int j = 0; j < nh; j++) {
        if ((best_sel >> j) & 1U) {
            hb.push_back(ps.hb[j]);
        }
    }
    for (int j = 0; j < nh; j++) {
        if (!((best_sel >> j) & 1U)) {
            hb.push_back(ps.hb[j]);
        }
    }
    return hb;
}

template <class T>
void tile_program(int n, const TapeView &t, const TilePass &ps, int low, const std::vector<int> &woff,
                  const std::vector<int> &poff, std::vector<TrInst<T>> &prog, std::vector<TpOp<T>> &lops) {
    const int k = low + static_cast<int>(ps.hb.size());
    const uint32_t lane_bits = 63U, reg_bits = ((1U << k) - 1) & ~lane_bits;
    auto cx = [](const std::complex<T> &z) { return dcx<T>{z.real(), z.imag()}; };
    for (const auto &g : tile_ops_merged<T>(n, t, ps, low, woff, poff)) {
        TrInst<T> in{};
        if (g.kind == GroupOp<T>::Parity) { // parity of the (lane, register) parts of the mask
            const uint32_t pm = static_cast<uint32_t>(g.parity_mask);
            in.code = kTrParity;
            in.tmask = pm & lane_bits;
            in.rmask = pm >> 6;
            in.mat[0] = cx(g.parity_phase[0]);
            in.mat[1] = cx(g.parity_phase[1]);
            prog.push_back(in);
            continue;
        }
        uint32_t pins = 0, xo = 0;
        for (int b = 0; b < g.nb; b++) {
            pins |= 1U << g.pos[b];
        }
        for (int b = 0; b < g.M; b++) {
            xo |= static_cast<uint32_t>(g.off[b]);
        }
        const uint32_t fx = static_cast<uint32_t>(g.fixed);
        const bool clean = (fx & ~pins) == 0 && (fx & xo) == 0 && (xo & ~pins) == 0;
        // condition on the non-member pins, split into lane and register bits
        const uint32_t cmask = pins & ~xo, cval = fx & cmask;
        in.tmask = cmask & lane_bits;
        in.tval = cval & lane_bits;
        in.rmask = cmask >> 6;
        in.rval = cval >> 6;
        const int nx = __builtin_popcount(xo);
        if (clean && (g.M == 1 || g.diag) && g.M <= 4 && nx <= 2) {
            in.code = kTrDiag;
            // phase index c: bit i = index bit mb[i]; mat[c] = the phase there
            // (1: no member, or the condition fails).  With at most two
            // pinned bits in all, the condition bits join the index and the
            // condition disappears.
            const bool fold = __builtin_popcount(pins) <= 2;
            const uint32_t ix = fold ? pins : xo;
            if (fold) {
                in.tmask = in.tval = in.rmask = in.rval = 0;
            }
            int mb[2] = {-1, -1}, nm = 0;
            for (int b = 0; b < k; b++) {
                if ((ix >> b) & 1U) {
                    mb[nm++] = b;
                }
            }
            for (int c = 0; c < 4; c++) {
                uint32_t u = 0;
                for (int a = 0; a < nm; a++) {
                    if ((c >> a) & 1) {
                        u |= 1U << mb[a];
                    }
                }
                in.mat[c] = dcx<T>{T(1), T(0)};
                if (c >= (1 << nm) || (fold && (u & cmask) != cval)) {
                    continue;
                }
                const uint32_t x = u & xo;
                for (int j = 0; j < g.M; j++) {
                    if (static_cast<uint32_t>(g.off[j]) == x) {
                        in.mat[c] = cx(g.mat[j * g.M + j]);
                    }
                }
            }
            // member bits held in registers feed slots ordered by register bit
            in.tcb[0] = in.tcb[1] = -1;
            in.nr = 0;
            for (int a = 0; a < nm; a++) {
                if (mb[a] >= 6) {
                    (in.nr == 0 ? in.ja : in.jb) = mb[a] - 6;
                    in.rcb[in.nr++] = a;
                } else {
                    in.tcb[a] = mb[a];
                }
            }
            prog.push_back(in);
            continue;
        }
        const bool p1 = clean && g.M == 2 && g.off[0] == 0 && nx == 1;
        const bool p2 = clean && g.M == 2 && nx == 2 && __builtin_popcountll(g.off[0]) == 1 &&
                        __builtin_popcountll(g.off[1]) == 1;
        const bool p4 = clean && g.M == 4 && g.off[0] == 0 && nx == 2 && __builtin_popcountll(g.off[1]) == 1 &&
                        __builtin_popcountll(g.off[2]) == 1 && g.off[3] == (g.off[1] | g.off[2]);
        // members in registers (a 4x4 op conditioned on lane bits goes through LDS)
        if ((p1 || p2 || (p4 && in.tmask == 0)) && (xo & ~reg_bits) == 0) {
            std::vector<std::complex<T>> m = g.mat;
            if (p1) {
                in.code = kTrP1;
                in.ja = __builtin_ctz(xo) - 6;
            } else {
                in.code = p2 ? kTrP2 : kTrP4;
                const int ra = __builtin_ctzll(g.off[p2 ? 0 : 1]) - 6;
                const int rb = __builtin_ctzll(g.off[p2 ? 1 : 2]) - 6;
                in.ja = std::min(ra, rb);
                in.jb = std::max(ra, rb);
                if (ra > rb) { // members listed in the other order: permute the matrix
                    const int M = g.M;
                    const std::vector<int> perm = M == 2 ? std::vector<int>{1, 0} : std::vector<int>{0, 2, 1, 3};
                    for (int r = 0; r < M; r++) {
                        for (int c = 0; c < M; c++) {
                            m[r * M + c] = g.mat[perm[r] * M + perm[c]];
                        }
                    }
                }
            }
            for (int b = 0; b < g.M * g.M; b++) {
                in.mat[b] = cx(m[b]);
            }
            prog.push_back(in);
            continue;
        }
        if (p1) { // 2x2 on a lane bit
            in.code = kTrLane;
            in.lx = xo;
            for (int b = 0; b < 4; b++) {
                in.mat[b] = cx(g.mat[b]);
            }
            prog.push_back(in);
            continue;
        }
        if (clean && g.kind == GroupOp<T>::Group && g.M <= 4 && nx == 2 && __builtin_ctz(xo) < 6) {
            // two member bits, at least one a lane bit: the op as a 4x4 matrix U
            // over the patterns u of (mb[0], mb[1]) (identity where no member)
            int mb[2], nm = 0;
            for (int b = 0; b < k; b++) {
                if ((xo >> b) & 1U) {
                    mb[nm++] = b;
                }
            }
            auto pat = [&](uint64_t off) { return static_cast<int>(((off >> mb[0]) & 1U) | (((off >> mb[1]) & 1U) << 1)); };
            std::complex<T> U[4][4] = {};
            for (int a = 0; a < 4; a++) {
                U[a][a] = 1;
            }
            for (int a = 0; a < g.M; a++) {
                for (int c = 0; c < g.M; c++) {
                    U[pat(g.off[a])][pat(g.off[c])] = g.mat[a * g.M + c];
                }
            }
            if (mb[1] >= 6) { // mb[0] lane bit, mb[1] register bit
                const int iq = 0, ij = 1;
                in.code = kTrMix1;
                in.lx = 1U << mb[iq];
                in.ja = mb[ij] - 6;
                auto u = [&](int qb, int jb) { return (qb << iq) | (jb << ij); };
                for (int oq = 0; oq < 2; oq++) {
                    for (int jb = 0; jb < 2; jb++) {
                        const int row = u(oq, jb);
                        const int cols[4] = {u(oq, 0), u(oq, 1), u(1 - oq, 0), u(1 - oq, 1)};
                        for (int c = 0; c < 4; c++) {
                            in.mat[8 * oq + 4 * jb + c] = cx(U[row][cols[c]]);
                        }
                    }
                }
            } else { // both lane bits
                in.code = kTrMix2;
                in.lx = 1U << mb[0];
                in.lx2 = 1U << mb[1];
                for (int o = 0; o < 4; o++) {
                    for (int d = 0; d < 4; d++) {
                        in.mat[4 * o + d] = cx(U[o][o ^ d]);
                    }
                }
            }
            prog.push_back(in);
            continue;
        }
        in = TrInst<T>{};
        in.code = kTrLds;
        in.lds_op = static_cast<int>(lops.size());
        lops.push_back(tile_lds_op(g));
        prog.push_back(in);
    }
}

} // namespace plamd
