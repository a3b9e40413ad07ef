// C-ABI of the gfx950 state-vector engine (declared in include/pl_amd.h):
// state lifecycle, gates, measurements, sampling, the adjoint and VJP sweeps,
// the fused and tiled executors.  The batched-circuit and MLP entry points
// are in batch.hip and mlp.hip.
//
// Host orchestration: every entry point validates its arguments, lowers the
// request (lower.hpp) and enqueues kernels (kernels_*.hpp) on the handle's
// HIP stream.  There is no CPU compute path: if the device work cannot be
// enqueued the call fails with PL_ERR_DEVICE.
#include "engine_common.hpp"
#include "kernels_adjoint.hpp"
#include "kernels_reduce.hpp"
#include "kernels_sample.hpp"
#include "kernels_mfma.hpp"
#include "kernels_mcmc.hpp"
#include "fuse.hpp"
#include "tile.hpp"
#include "tiled_launch.hpp"

namespace plamd {

// Copies between a caller's (pageable) host array and device memory, on the
// handle's stream.  Below kPinMin bytes each is one hipMemcpyAsync.  Above,
// chunks of kPinSlot bytes alternate between two pinned slots: up to
// kPinThreads host threads copy one chunk between the caller's array and a
// slot while the DMA engine moves the neighbouring chunk through the other
// slot.  The slots (2 x 64 MiB) belong to the process: allocated on first
// use, kept until it ends, and taken by one transfer at a time (every handle
// is on device 0).  Where they cannot be pinned every copy is a plain one.
constexpr size_t kPinSlot = size_t{64} << 20;
constexpr size_t kPinMin = size_t{8} << 20;
constexpr unsigned kPinThreads = 8;

struct PinRing {
    std::mutex mu;       // held for a whole transfer
    void *base = nullptr;
    bool unavailable = false; // pinning was refused once: plain copies from then on
    hipEvent_t ev[2] = {nullptr, nullptr}; // recorded after the last copy queued through a slot
    bool busy[2] = {false, false};         // such a copy exists: wait on ev before reuse
};
static PinRing g_pin;

// Allocates the slots on first use; false if the process cannot pin them
// (a locked-memory limit, say), and the transfers then use plain copies.
static bool pin_ready() {
    if (!g_pin.base && !g_pin.unavailable) {
        for (hipEvent_t &e : g_pin.ev) {
            if (!e) {
                HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            }
        }
        if (hipHostMalloc(&g_pin.base, 2 * kPinSlot, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError(); // the refusal is not left for the next call to report
            g_pin.base = nullptr;
            g_pin.unavailable = true;
        }
    }
    return g_pin.base != nullptr;
}

static char *pin_slot(int s) {
    return static_cast<char *>(g_pin.base) + s * kPinSlot;
}

// Waits until no queued copy reads or writes slot s.
static void pin_wait(int s) {
    if (g_pin.busy[s]) {
        HIP_OK(hipEventSynchronize(g_pin.ev[s]));
        g_pin.busy[s] = false;
    }
}

// Queues a copy through slot s on h's stream and marks the slot busy behind it.
static void pin_copy(pl_sv_t h, int s, void *dst, const void *src, size_t len, hipMemcpyKind kind) {
    HIP_OK(hipMemcpyAsync(dst, src, len, kind, h->stream));
    HIP_OK(hipEventRecord(g_pin.ev[s], h->stream));
    g_pin.busy[s] = true;
}

// memcpy of len bytes split into page-aligned parts, one per thread (the
// caller's thread takes the first)
static void parallel_memcpy(char *dst, const char *src, size_t len) {
    const unsigned nthr = std::min(kPinThreads, std::max(1U, std::thread::hardware_concurrency()));
    const size_t per = ((len + nthr - 1) / nthr + 4095) & ~size_t{4095};
    std::vector<std::thread> th;
    auto join_all = [&] {
        for (auto &t : th) {
            t.join();
        }
    };
    try {
        for (size_t p = per; p < len; p += per) {
            th.emplace_back([=] { std::memcpy(dst + p, src + p, std::min(per, len - p)); });
        }
    } catch (...) {
        join_all(); // a thread that could not start: finish the ones that did, then fail
        throw;
    }
    std::memcpy(dst, src, std::min(per, len));
    join_all();
}

// Host to device: chunk c is copied into slot c % 2 once that slot's previous
// copy has left it.  The copies stay queued when this returns; the caller
// keeps src alive until it synchronises the stream.
static void upload_host(pl_sv_t h, void *dst, const void *src, size_t bytes) {
    std::unique_lock<std::mutex> lock(g_pin.mu, std::defer_lock);
    if (bytes >= kPinMin) {
        lock.lock();
    }
    if (bytes < kPinMin || !pin_ready()) {
        HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
        return;
    }
    auto *out = static_cast<char *>(dst);
    const auto *in = static_cast<const char *>(src);
    for (size_t off = 0, c = 0; off < bytes; off += kPinSlot, c++) {
        const size_t len = std::min(kPinSlot, bytes - off);
        const int s = static_cast<int>(c & 1);
        char *slot = pin_slot(s);
        pin_wait(s);
        parallel_memcpy(slot, in + off, len);
        pin_copy(h, s, out + off, slot, len, hipMemcpyHostToDevice);
    }
}

// Device to host, after the work queued before it: the DMA of chunk c into
// slot c % 2 is queued before chunk c - 1 is copied out of the other slot.
// Returns when every byte is in dst.
static void download_host(pl_sv_t h, void *dst, const void *src, size_t bytes) {
    std::unique_lock<std::mutex> lock(g_pin.mu, std::defer_lock);
    if (bytes >= kPinMin) {
        lock.lock();
    }
    if (bytes < kPinMin || !pin_ready()) {
        HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        return;
    }
    auto *out = static_cast<char *>(dst);
    const auto *in = static_cast<const char *>(src);
    const size_t nchunks = (bytes + kPinSlot - 1) / kPinSlot;
    auto drain = [&](size_t c) { // chunk c from its slot to dst
        const int s = static_cast<int>(c & 1);
        const size_t off = c * kPinSlot, len = std::min(kPinSlot, bytes - off);
        pin_wait(s);
        parallel_memcpy(out + off, pin_slot(s), len);
    };
    for (size_t c = 0; c < nchunks; c++) {
        const int s = static_cast<int>(c & 1);
        const size_t off = c * kPinSlot, len = std::min(kPinSlot, bytes - off);
        char *slot = pin_slot(s);
        // slot s last held chunk c - 2, copied out by the drain below; the wait
        // covers a copy through it from an earlier transfer, on any handle's stream
        pin_wait(s);
        pin_copy(h, s, slot, in + off, len, hipMemcpyDeviceToHost);
        if (c > 0) {
            drain(c - 1);
        }
    }
    drain(nchunks - 1);
}

template <class T> struct Engine {
    using V = v2<T>;
    using Scalar = T;
    pl_sv *h;
    explicit Engine(pl_sv *hh) : h(hh) {}
    V *psi() const { return static_cast<V *>(h->d); }
    uint64_t len() const { return h->len; }
    int n() const { return h->n; }

    template <int M> void launch_small(const GroupOp<T> &g, V *sv) {
        SmallOp<T, M> op{};
        op.geo.fixed = g.fixed;
        op.geo.nb = g.nb;
        for (int i = 0; i < g.nb; i++) {
            op.geo.pos[i] = g.pos[i];
        }
        for (int j = 0; j < M; j++) {
            op.off[j] = g.off[j];
        }
        for (int j = 0; j < M * M; j++) {
            op.mat[j] = {g.mat[j].real(), g.mat[j].imag()};
        }
        op.diag = g.diag ? 1 : 0;
        const uint64_t nwork = len() >> g.nb;
        if constexpr (M == 2) {
            // target bit inside a wavefront's 64 consecutive amplitudes and
            // no other pinned bit there: lane-per-amplitude shuffle kernel
            const uint64_t t = g.off[1];
            const bool single = g.off[0] == 0 && t && !(t & (t - 1));
            const int p = single ? __builtin_ctzll(t) : 64;
            bool lane_ok = p < 6 && (nwork << 1) >= 64;
            for (int i = 0; i < g.nb && lane_ok; i++) {
                lane_ok = g.pos[i] == p || g.pos[i] >= 6;
            }
            if (lane_ok) {
                Geom geo{};
                geo.fixed = g.fixed;
                for (int i = 0; i < g.nb; i++) {
                    if (g.pos[i] != p) {
                        geo.pos[geo.nb++] = g.pos[i];
                    }
                }
                const uint64_t nelem = nwork << 1;
                const dim3 grid(static_cast<unsigned>(nelem / kBlock + (nelem % kBlock ? 1 : 0)));
                if (g.diag) {
                    hipLaunchKernelGGL((k_pair_lane<T, true>), grid, dim3(kBlock), 0, h->stream, sv, nelem, geo, p,
                                       op.mat[0], op.mat[1], op.mat[2], op.mat[3]);
                } else {
                    hipLaunchKernelGGL((k_pair_lane<T, false>), grid, dim3(kBlock), 0, h->stream, sv, nelem, geo,
                                       p, op.mat[0], op.mat[1], op.mat[2], op.mat[3]);
                }
                HIP_OK(hipGetLastError());
                return;
            }
        }
        // a pinned bit inside a 128-byte line: stage whole lines through LDS
        constexpr int line_bits = sizeof(T) == 8 ? 3 : 4;
        int nhi = 0;
        for (int i = 0; i < g.nb; i++) {
            nhi += g.pos[i] >= kTileLow ? 1 : 0;
        }
        if (g.nb > 0 && g.pos[0] < line_bits && nhi <= kTileMaxHi && n() >= kTileLow + nhi) {
            TileOp<T, M> to{};
            uint32_t hmap_pos[kTileMaxHi] = {};
            for (int i = 0; i < g.nb; i++) {
                if (g.pos[i] >= kTileLow) {
                    hmap_pos[to.hi.nb] = g.pos[i];
                    to.hi.pos[to.hi.nb++] = g.pos[i];
                } else {
                    to.lo.pos[to.lo.nb++] = g.pos[i];
                }
            }
            auto compress_h = [&](uint64_t bits) {
                uint32_t hbits = 0;
                for (int k = 0; k < to.hi.nb; k++) {
                    if ((bits >> hmap_pos[k]) & 1U) {
                        hbits |= 1U << k;
                    }
                }
                return hbits;
            };
            constexpr uint64_t low_mask = (uint64_t{1} << kTileLow) - 1;
            to.fixed_lo = static_cast<uint32_t>(g.fixed & low_mask);
            to.fixed_h = compress_h(g.fixed);
            for (int j = 0; j < M; j++) {
                to.off_e[j] = static_cast<uint32_t>(g.off[j] & low_mask);
                to.off_h[j] = compress_h(g.off[j]);
                for (int c = 0; c < M; c++) {
                    to.mat[j * M + c] = op.mat[j * M + c];
                }
            }
            to.diag = op.diag;
            const uint64_t ntiles = len() >> (kTileLow + nhi);
            const unsigned grid = static_cast<unsigned>(ntiles);
            if (g.diag) {
                launch_tile<M, true>(nhi, grid, sv, ntiles, to);
            } else {
                launch_tile<M, false>(nhi, grid, sv, ntiles, to);
            }
            HIP_OK(hipGetLastError());
            return;
        }
        const unsigned grid = static_cast<unsigned>((nwork + kBlock - 1) / kBlock);
        if (g.diag) {
            hipLaunchKernelGGL((k_group_flat<T, M, true>), dim3(grid), dim3(kBlock), 0, h->stream, sv, nwork, op);
        } else {
            hipLaunchKernelGGL((k_group_flat<T, M, false>), dim3(grid), dim3(kBlock), 0, h->stream, sv, nwork, op);
        }
        HIP_OK(hipGetLastError());
    }

    template <int M, bool DIAG>
    void launch_tile(int nhi, unsigned grid, V *sv, uint64_t ntiles, const TileOp<T, M> &to) {
        switch (nhi) {
        case 0: hipLaunchKernelGGL((k_group_tile<T, M, 0, DIAG>), dim3(grid), dim3(kBlock), 0, h->stream, sv, ntiles, to); break;
        case 1: hipLaunchKernelGGL((k_group_tile<T, M, 1, DIAG>), dim3(grid), dim3(kBlock), 0, h->stream, sv, ntiles, to); break;
        case 2: hipLaunchKernelGGL((k_group_tile<T, M, 2, DIAG>), dim3(grid), dim3(kBlock), 0, h->stream, sv, ntiles, to); break;
        default: hipLaunchKernelGGL((k_group_tile<T, M, 3, DIAG>), dim3(grid), dim3(kBlock), 0, h->stream, sv, ntiles, to); break;
        }
    }

    template <int M> void launch_big(const GroupOp<T> &g, V *sv) {
        std::vector<dcx<T>> m(M * M);
        for (int j = 0; j < M * M; j++) {
            m[j] = {g.mat[j].real(), g.mat[j].imag()};
        }
        launch_big_dev<M>(g, sv, static_cast<const dcx<T> *>(stage(h, m.data(), m.size() * sizeof(dcx<T>))));
        HIP_OK(hipStreamSynchronize(h->stream)); // staging is reused by the next call
    }

    // dmat: the M x M matrix already in device memory
    template <int M> void launch_big_dev(const GroupOp<T> &g, V *sv, const dcx<T> *dmat) {
        BigOp<T, M> op{};
        op.geo.fixed = g.fixed;
        op.geo.nb = g.nb;
        for (int i = 0; i < g.nb; i++) {
            op.geo.pos[i] = g.pos[i];
        }
        for (int j = 0; j < M; j++) {
            op.off[j] = g.off[j];
        }
        op.mat = dmat;
        const uint64_t nwork = len() >> g.nb;
        hipLaunchKernelGGL((k_group_big<T, M>), dim3(grid_for(nwork)), dim3(kBlock), 0, h->stream,
                           sv, nwork, op);
        HIP_OK(hipGetLastError());
    }

    void launch_huge(const GroupOp<T> &g, V *sv) {
        const size_t M = g.off.size();
        std::vector<unsigned char> blob(M * 8 + M * M * sizeof(dcx<T>));
        std::memcpy(blob.data(), g.off.data(), M * 8);
        auto *m = reinterpret_cast<dcx<T> *>(blob.data() + M * 8);
        for (size_t j = 0; j < M * M; j++) {
            m[j] = {g.mat[j].real(), g.mat[j].imag()};
        }
        auto *dev = static_cast<unsigned char *>(stage(h, blob.data(), blob.size()));
        Geom geo{};
        geo.fixed = g.fixed;
        geo.nb = g.nb;
        for (int i = 0; i < g.nb; i++) {
            geo.pos[i] = g.pos[i];
        }
        const uint64_t nwork = len() >> g.nb;
        const unsigned grid = static_cast<unsigned>(std::min<uint64_t>(nwork, device_cus() * 8));
        hipLaunchKernelGGL((k_group_huge<T>), dim3(grid), dim3(kBlock), M * sizeof(V), h->stream,
                           sv, nwork, geo, static_cast<int>(M),
                           reinterpret_cast<const uint64_t *>(dev),
                           reinterpret_cast<const dcx<T> *>(dev + M * 8));
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(h->stream));
    }

    void apply(const GroupOp<T> &g, V *sv) {
        if (g.kind == GroupOp<T>::Noop) {
            return;
        }
        if (g.kind == GroupOp<T>::Parity) {
            hipLaunchKernelGGL((k_parity<T>), dim3(grid_for(len())), dim3(kBlock), 0, h->stream, sv,
                               len(), g.parity_mask,
                               dcx<T>{g.parity_phase[0].real(), g.parity_phase[0].imag()},
                               dcx<T>{g.parity_phase[1].real(), g.parity_phase[1].imag()});
            HIP_OK(hipGetLastError());
            return;
        }
        switch (g.M) {
        case 1: return launch_small<1>(g, sv);
        case 2: return launch_small<2>(g, sv);
        case 4: return launch_small<4>(g, sv);
        case 8: return launch_big<8>(g, sv);
        case 16: return launch_big<16>(g, sv);
        case 32: return launch_big<32>(g, sv);
        default: return launch_huge(g, sv);
        }
    }

    std::vector<T> params_t(const double *p, int np) const {
        std::vector<T> v(np);
        for (int i = 0; i < np; i++) {
            v[i] = static_cast<T>(p[i]);
        }
        return v;
    }

    void gate(V *sv, int id, const int *w, int nw, bool inv, const double *p, int np) {
        const auto pt = params_t(p, np);
        apply(lower_gate<T>(n(), id, w, nw, inv, pt.data(), np), sv);
    }

    void matrix(V *sv, const double *m, const int *w, int nw, bool inv) {
        if (nw <= 0) {
            throw InvalidArg("Number of wires must be larger than 0");
        }
        if (nw > kMaxPos) {
            throw InvalidArg("matrices on more than 12 wires are not supported");
        }
        const size_t dim = size_t{1} << (2 * nw);
        std::vector<std::complex<T>> mt(dim);
        for (size_t i = 0; i < dim; i++) {
            mt[i] = {static_cast<T>(m[2 * i]), static_cast<T>(m[2 * i + 1])};
        }
        apply(lower_matrix<T>(n(), mt.data(), w, nw, inv), sv);
    }

    double generator(V *sv, int id, const int *w, int nw) {
        T s = 0;
        const auto g = lower_generator<T>(n(), id, w, nw, false, &s);
        apply(g, sv);
        return static_cast<double>(s);
    }

    // ---- precision path: fused blocks -----------------------------------------
    static constexpr bool kMfma = std::is_same_v<T, float>;
    bool mfma_ok(const FusedBlock &b, int flags) const {
        return kMfma && !(flags & PL_FUSE_NO_MFMA) && n() >= 5 + kF5ColBits && b.fused && b.bits.size() == 5;
    }

    // tile geometry of k_fused5_mfma: 5 target bits + the kF5ColBits lowest other bits
    Fused5Geom fused5_geom(const std::vector<int> &tb) const {
        std::vector<std::pair<int, uint32_t>> all;
        for (int m = 0; m < 5; m++) {
            all.emplace_back(tb[m], (1U << kF5ColBits) << m);
        }
        for (int p = 0, m = 0; m < kF5ColBits; p++) {
            if (std::find(tb.begin(), tb.end(), p) == tb.end()) {
                all.emplace_back(p, 1U << m);
                m++;
            }
        }
        std::sort(all.begin(), all.end());
        Fused5Geom g{};
        for (int k = 0; k < 5 + kF5ColBits; k++) {
            g.pos[k] = all[k].first;
            g.lds_of[k] = all[k].second;
        }
        return g;
    }

    // Applies a tape as fused blocks; returns the number of blocks.  Each
    // block is planned, built, uploaded and launched before the next is
    // planned, so the host's fusion overlaps the device's blocks.
    int apply_fused(V *sv, const TapeView &t, int max_wires, int flags, std::vector<FusedBlock> *keep) {
        const bool probe_mfma = kMfma && !(flags & PL_FUSE_NO_MFMA) && n() >= 5 + kF5ColBits;
        FuseScheduler sched(n(), t, max_wires, probe_mfma ? 5 : 0);
        const bool renorm = flags & PL_FUSE_RENORM;

        // block matrices go through kFuseSlots slots of host and device memory;
        // when they run out, the blocks queued so far are waited for and the
        // slots start over, so a host slot is never rewritten while its copy
        // may be pending and a device slot never while its block may run
        constexpr size_t kFuseSlots = 256;
        const size_t slot = (std::max(sizeof(Fused5Mat), size_t{32 * 32} * sizeof(dcx<T>)) + 255) & ~size_t{255};
        const uint64_t ntiles = n() >= 5 + kF5ColBits ? (len() >> (5 + kF5ColBits)) : 1;
        const size_t mbytes = kFuseSlots * slot;
        std::vector<unsigned char> host(mbytes);
        unsigned char *dev = nullptr;
        HIP_OK(hipMalloc(&dev, mbytes + ntiles * sizeof(double) + 256));
        struct Free {
            unsigned char *p;
            hipStream_t s;
            ~Free() {
                (void)hipStreamSynchronize(s);
                (void)hipFree(p);
            }
        } guard{dev, h->stream};
        auto *partial = reinterpret_cast<double *>(dev + mbytes);
        auto *scale = partial + ntiles;
        bool pending = false; // scale holds a factor not yet applied

        std::vector<int> woff(t.nops + 1, 0), poff(t.nops + 1, 0);
        for (int i = 0; i < t.nops; i++) {
            woff[i + 1] = woff[i] + t.nwires[i];
            poff[i + 1] = poff[i] + t.nparams[i];
        }
        size_t used_slots = 0;
        int nb = 0;
        FusedBlock b;
        while (sched.next(b)) {
            nb++;
            if (!b.fused) {
                const int o = b.ops[0];
                gate(sv, t.ids[o], t.wires + woff[o], t.nwires[o], t.inverse[o] != 0, t.params + poff[o],
                     t.nparams[o]);
            } else {
                const bool mf = mfma_ok(b, flags);
                const bool big = !mf && b.U.size() >= 64; // matrix in device memory
                unsigned char *hs = nullptr, *ds = nullptr;
                if (mf || big) {
                    if (used_slots == kFuseSlots) {
                        HIP_OK(hipStreamSynchronize(h->stream));
                        used_slots = 0;
                    }
                    hs = host.data() + used_slots * slot;
                    ds = dev + used_slots * slot;
                    used_slots++;
                }
                if (mf) {
                    auto *fm = reinterpret_cast<Fused5Mat *>(hs);
                    for (int l = 0; l < 64; l++) {
                        for (int st = 0; st < 16; st++) {
                            const auto u = b.U[(l & 31) * 32 + 2 * st + (l >> 5)];
                            fm->a[l][st] = float2{static_cast<float>(u.real()), static_cast<float>(u.imag())};
                        }
                    }
                    HIP_OK(hipMemcpyAsync(ds, hs, sizeof(Fused5Mat), hipMemcpyHostToDevice, h->stream));
                    if constexpr (kMfma) {
                        // persistent grid: the three workgroups a CU holds at once, contiguous tile ranges
                        const uint64_t want = std::min<uint64_t>(ntiles, uint64_t(device_cus()) * 3);
                        const uint64_t per = (ntiles + want - 1) / want;
                        const uint64_t nblk = (ntiles + per - 1) / per; // no empty workgroup
                        hipLaunchKernelGGL(k_fused5_mfma<kF5ColBits>, dim3(static_cast<unsigned>(nblk)),
                                           dim3(2 << kF5ColBits), 0, h->stream, sv, ntiles, per,
                                           fused5_geom(b.bits), reinterpret_cast<const Fused5Mat *>(ds),
                                           pending ? scale : nullptr, renorm ? partial : nullptr);
                        HIP_OK(hipGetLastError());
                        if (renorm) {
                            hipLaunchKernelGGL(k_norm_scale, dim3(1), dim3(1024), 0, h->stream, partial, nblk,
                                               scale);
                            HIP_OK(hipGetLastError());
                            pending = true;
                        } else {
                            pending = false;
                        }
                    }
                } else {
                    std::vector<std::complex<T>> mt(b.U.size());
                    for (size_t j = 0; j < b.U.size(); j++) {
                        mt[j] = {static_cast<T>(b.U[j].real()), static_cast<T>(b.U[j].imag())};
                    }
                    if (big) {
                        auto *m = reinterpret_cast<dcx<T> *>(hs);
                        for (size_t j = 0; j < b.U.size(); j++) {
                            m[j] = {mt[j].real(), mt[j].imag()};
                        }
                        HIP_OK(hipMemcpyAsync(ds, hs, b.U.size() * sizeof(dcx<T>), hipMemcpyHostToDevice,
                                              h->stream));
                    }
                    const auto g = make_group<T>(b.bits, 0, block_offsets(b.bits), std::move(mt));
                    switch (g.M) {
                    case 8: launch_big_dev<8>(g, sv, reinterpret_cast<const dcx<T> *>(ds)); break;
                    case 16: launch_big_dev<16>(g, sv, reinterpret_cast<const dcx<T> *>(ds)); break;
                    case 32: launch_big_dev<32>(g, sv, reinterpret_cast<const dcx<T> *>(ds)); break;
                    default: apply(g, sv); break; // M <= 4: matrix in the kernel arguments
                    }
                }
            }
            if (keep) {
                keep->push_back(std::move(b));
            }
        }
        if (renorm) {
            if (!pending) { // no MFMA block produced a norm: one FP64 norm pass
                const double nrm = dot(sv, sv).real();
                const double sc = nrm > 0 ? 1.0 / std::sqrt(nrm) : 1.0;
                HIP_OK(hipMemcpyAsync(scale, &sc, sizeof(double), hipMemcpyHostToDevice, h->stream));
                HIP_OK(hipStreamSynchronize(h->stream));
            }
            hipLaunchKernelGGL((k_scale<T>), dim3(grid_for(len())), dim3(kBlock), 0, h->stream, sv, len(),
                               scale);
            HIP_OK(hipGetLastError());
        }
        HIP_OK(hipStreamSynchronize(h->stream));
        return nb;
    }

    // psi <- psi / ||psi|| with the norm accumulated in FP64 (one read pass and
    // one scaling pass)
    void renormalize(V *sv) {
        const double nrm = dot(sv, sv).real();
        const double sc = nrm > 0 ? 1.0 / std::sqrt(nrm) : 1.0;
        auto *d = static_cast<double *>(buf(h, kBufScalar, sizeof(double)));
        HIP_OK(hipMemcpyAsync(d, &sc, sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        hipLaunchKernelGGL((k_scale<T>), dim3(grid_for(len())), dim3(kBlock), 0, h->stream, sv, len(), d);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(h->stream));
    }

    // ---- tile passes: many gates per read and write of the state --------------
    // defaults measured on the 50-layer random circuit (30 qubits complex128,
    // 33 qubits complex64): the largest tile each precision's registers hold
    // wins (fewer passes outweigh the lower occupancy)
    static constexpr int kTileDefaultBits = std::is_same_v<T, double> ? 11 : 12;
    static constexpr int kTileMaxBitsC128 = 11;
    static constexpr int kTileDefaultLow = std::is_same_v<T, double> ? 3 : 4;

    // Applies a tape as tile passes (tile.hpp); returns the number of launches.
    // k = 0 / low < 0 pick the measured defaults of the precision; with k = 0
    // a state narrower than the smallest tile runs gate by gate.  Each pass is
    // planned, compiled and launched before the next is planned, so the host's
    // planning overlaps the device's passes.
    int apply_tiled(V *sv, const TapeView &t, int k, int low) {
        std::vector<int> woff(t.nops + 1, 0), poff(t.nops + 1, 0);
        for (int i = 0; i < t.nops; i++) {
            woff[i + 1] = woff[i] + t.nwires[i];
            poff[i + 1] = poff[i] + t.nparams[i];
        }
        if (k == 0) {
            if (n() < kTpMinBits) {
                int launches = 0;
                for (int i = 0; i < t.nops; i++) {
                    if (t.ids[i] != -1) {
                        gate(sv, t.ids[i], t.wires + woff[i], t.nwires[i], t.inverse[i] != 0, t.params + poff[i],
                             t.nparams[i]);
                        launches++;
                    }
                }
                HIP_OK(hipStreamSynchronize(h->stream));
                return launches;
            }
            k = std::min(n(), kTileDefaultBits);
        }
        if (low < 0) {
            low = std::min(kTileDefaultLow, k - 6);
        }
        TileScheduler sched(n(), t, k, low);
        if (std::is_same_v<T, double> && k > kTileMaxBitsC128) {
            throw InvalidArg("complex128 tiles hold at most 2^" + std::to_string(kTileMaxBitsC128) +
                             " amplitudes (a wavefront's registers)");
        }
        int dev_id = 0, lds_max = 0;
        HIP_OK(hipGetDevice(&dev_id));
        HIP_OK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_id));
        if ((size_t{1} << k) * sizeof(V) > static_cast<size_t>(lds_max)) {
            throw InvalidArg("a tile of 2^" + std::to_string(k) + " amplitudes exceeds the LDS of a workgroup");
        }
        // the passes' programs, uploaded pass by pass into one device buffer
        // (instructions, then the LDS-applied ops).  A pass's program holds at
        // most one instruction and one LDS op per operation.  The buffers hold
        // `cap` of each; when the next pass could overflow them, the passes
        // queued so far are waited for and the buffers start over (grown if
        // that one pass needs more), so host copies never move while a copy
        // may read them and no queued pass reads a slot being rewritten.
        std::vector<TrInst<T>> prog;
        std::vector<TpOp<T>> lops;
        size_t cap = 0, pbytes = 0;
        struct Free {
            void *p = nullptr;
            hipStream_t s;
            ~Free() {
                if (p) {
                    (void)hipStreamSynchronize(s);
                    (void)hipFree(p);
                }
            }
        } guard{nullptr, h->stream};
        TrInst<T> *dprog = nullptr;
        TpOp<T> *dlops = nullptr;
        auto room_for = [&](size_t nops_pass) {
            if (prog.size() + nops_pass <= cap && lops.size() + nops_pass <= cap) {
                return;
            }
            HIP_OK(hipStreamSynchronize(h->stream));
            prog.clear();
            lops.clear();
            if (nops_pass > cap) {
                if (guard.p) {
                    HIP_OK(hipFree(guard.p));
                    guard.p = nullptr;
                }
                cap = std::max<size_t>(nops_pass, std::min<size_t>(static_cast<size_t>(t.nops), kTileLookahead));
                prog.reserve(cap);
                lops.reserve(cap);
                pbytes = (cap * sizeof(TrInst<T>) + 255) & ~size_t{255};
                HIP_OK(hipMalloc(&guard.p, pbytes + cap * sizeof(TpOp<T>)));
                dprog = reinterpret_cast<TrInst<T> *>(guard.p);
                dlops = reinterpret_cast<TpOp<T> *>(static_cast<unsigned char *>(guard.p) + pbytes);
            }
        };
        const uint64_t ntiles = uint64_t{1} << (n() - k);
        int npasses = 0;
        TilePass ps;
        while (sched.next(ps)) {
            npasses++;
            if (!ps.tiled) {
                const int o = ps.ops[0];
                gate(sv, t.ids[o], t.wires + woff[o], t.nwires[o], t.inverse[o] != 0, t.params + poff[o],
                     t.nparams[o]);
                continue;
            }
            ps.hb = tile_lane_order<T>(n(), t, ps, low, woff, poff); // lane bits where its ops are cheapest
            room_for(ps.ops.size());
            const size_t p0 = prog.size(), l0 = lops.size();
            tile_program<T>(n(), t, ps, low, woff, poff, prog, lops);
            if (prog.size() > cap || lops.size() > cap) {
                throw std::logic_error("a tile program longer than its pass");
            }
            const int ni = static_cast<int>(prog.size() - p0);
            if (ni == 0) {
                continue;
            }
            HIP_OK(hipMemcpyAsync(dprog + p0, prog.data() + p0, ni * sizeof(TrInst<T>), hipMemcpyHostToDevice,
                                  h->stream));
            if (lops.size() > l0) {
                HIP_OK(hipMemcpyAsync(dlops + l0, lops.data() + l0, (lops.size() - l0) * sizeof(TpOp<T>),
                                      hipMemcpyHostToDevice, h->stream));
            }
            TpGeom g{};
            g.k = k;
            g.low = low;
            g.outer.nb = static_cast<int>(ps.hb.size());
            g.outer.fixed = 0;
            std::vector<int> asc = ps.hb;
            std::sort(asc.begin(), asc.end());
            for (size_t j = 0; j < ps.hb.size(); j++) {
                g.hb[j] = ps.hb[j];
                g.outer.pos[j] = asc[j];
            }
            HIP_OK(launch_tile_pass<T>(k - 6, sv, ntiles, g, dprog + p0, ni, dlops, lops.size() > l0, h->stream));
        }
        HIP_OK(hipStreamSynchronize(h->stream));
        return npasses;
    }

    // ---- reductions ----------------------------------------------------------
    std::complex<double> dot(const V *a, const V *b) {
        const unsigned grid = std::min<unsigned>(grid_for(len(), 1, 4), kReduceBlocks);
        auto *part = reinterpret_cast<double2 *>(reduce_buf(h, (kReduceBlocks + 1) * sizeof(double2)));
        hipLaunchKernelGGL((k_dot<T>), dim3(grid), dim3(kBlock), 0, h->stream, a, b, len(), part);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, h->stream, part, (int)grid,
                           part + kReduceBlocks);
        HIP_OK(hipGetLastError());
        double2 r;
        HIP_OK(hipMemcpyAsync(&r, part + kReduceBlocks, sizeof(r), hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        return {r.x, r.y};
    }

    // <a|b> left in device memory at *slot, ordered on the engine's stream with
    // no wait; part holds kReduceBlocks partials and must not overlap slot.
    void dot_to(const V *a, const V *b, double2 *part, double2 *slot) {
        const unsigned grid = std::min<unsigned>(grid_for(len(), 1, 4), kReduceBlocks);
        hipLaunchKernelGGL((k_dot<T>), dim3(grid), dim3(kBlock), 0, h->stream, a, b, len(), part);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, h->stream, part, (int)grid, slot);
        HIP_OK(hipGetLastError());
    }

    // sum_t c_t <psi|P_t|psi> over Pauli words, grouped by x-mask
    // High bits per Pauli tile: 10 where a workgroup's LDS holds a tile of
    // 2^13 amplitudes (fewer passes), else 8.
    int pauli_high_bits() {
        int dev = 0, lds_max = 0;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t need = (size_t{1} << (3 + kPauliTileHighMax)) * sizeof(V) + 1024; // + block_sum2's arrays
        return static_cast<size_t>(lds_max) >= need ? kPauliTileHighMax : kPauliTileHigh;
    }

    // sum_t c_t <psi|P_t|psi> over Pauli words.  Distinct x-masks are packed
    // greedily into passes whose high bits fit one LDS tile (k_pauli_tile);
    // each pass reads the state once.  Masks wider than a tile get a pass of
    // their own (k_pauli_expval).
    // ext / ext_bytes: a caller-provided device area for the term tables and
    // partial sums (nothing is allocated then); nullptr: the handle's buffers.
    double pauli_expval(const V *sv, const std::vector<PauliWord> &words,
                        const std::vector<double> &coeffs, unsigned char *ext = nullptr,
                        size_t ext_bytes = 0) {
        std::map<uint64_t, std::vector<PauliTerm>> groups;
        for (size_t t = 0; t < words.size(); t++) {
            groups[words[t].x].push_back(PauliTerm{words[t].zy, coeffs[t], words[t].ny & 3, 0});
        }
        const int nlow = std::min(3, n());
        const uint64_t lowm = (uint64_t{1} << nlow) - 1;
        const int hmax = pauli_high_bits();
        // work items: an x-mask with all of its terms
        struct Item {
            uint64_t x;
            const std::vector<PauliTerm> *terms;
        };
        std::vector<Item> items;
        std::vector<uint64_t> wide;
        for (auto &kv : groups) {
            if (kv.first && __builtin_popcountll(kv.first & ~lowm) > hmax) {
                wide.push_back(kv.first);
                continue;
            }
            items.push_back(Item{kv.first, &kv.second});
        }
        // widest high parts first; the diagonal items (x = 0) fit any pass
        std::stable_sort(items.begin(), items.end(), [&](const Item &a, const Item &b) {
            return __builtin_popcountll(a.x & ~lowm) > __builtin_popcountll(b.x & ~lowm);
        });
        struct Pass {
            uint64_t high = 0;
            std::vector<const Item *> its;
        };
        std::vector<Pass> passes;
        std::vector<const Item *> rest;
        for (const auto &it : items) {
            rest.push_back(&it);
        }
        while (!rest.empty()) {
            Pass ps;
            std::vector<const Item *> left;
            for (const Item *it : rest) {
                const uint64_t xh = it->x & ~lowm;
                if (__builtin_popcountll(ps.high | xh) <= hmax) {
                    ps.high |= xh;
                    ps.its.push_back(it);
                } else {
                    left.push_back(it);
                }
            }
            passes.push_back(std::move(ps));
            rest.swap(left);
        }
        // every tile as large as the LDS budget allows: the high set is padded
        // with the lowest free positions (a pass whose masks need one high bit
        // would otherwise walk 2^(n-4) tiles of 16 amplitudes)
        for (auto &ps : passes) {
            for (int b = nlow; b < n() && __builtin_popcountll(ps.high) < hmax; b++) {
                ps.high |= uint64_t{1} << b;
            }
        }
        // per pass: geometry, groups and compressed terms, all staged together
        std::vector<PauliTileGeom> geos;
        std::vector<std::pair<size_t, size_t>> gspan; // offsets into the group table
        std::vector<std::pair<size_t, size_t>> tspan; // the pass's terms: first, count
        std::vector<PauliTileGroup> gtab;
        std::vector<PauliTileTerm> ttab;
        for (size_t pi = 0; pi < passes.size(); pi++) {
            const Pass &ps = passes[pi];
            PauliTileGeom g{};
            g.nlow = nlow;
            std::vector<int> pos;
            for (int b = 0; b < nlow; b++) {
                pos.push_back(b);
            }
            for (int b = 0; b < 64; b++) {
                if ((ps.high >> b) & 1) {
                    g.high[g.nhigh++] = b;
                    pos.push_back(b);
                }
            }
            const uint64_t smask = lowm | ps.high;
            auto compress = [&](uint64_t v) {
                uint32_t r = 0;
                for (size_t k = 0; k < pos.size(); k++) {
                    r |= static_cast<uint32_t>((v >> pos[k]) & 1) << k;
                }
                return r;
            };
            geos.push_back(g);
            const size_t g0 = gtab.size(), t_first = ttab.size();
            for (const Item *it : ps.its) {
                const uint64_t x = it->x;
                PauliTileGroup gr{};
                gr.xl = compress(x);
                const int nb = x ? kPauliPairs : 2 * kPauliPairs;
                std::vector<std::vector<PauliTileTerm>> bucket(nb);
                for (const auto &pt : *it->terms) {
                    const uint32_t zyl = compress(pt.zy & smask);
                    double cr = pt.coeff, ci = 0;
                    if (x != 0) { // see PauliTileTerm
                        const double sigma = (__builtin_popcount(gr.xl & zyl) & 1) ? -1.0 : 1.0;
                        const int ny = pt.ny & 3;
                        cr = pt.coeff * (1 + sigma) * (ny == 0 ? 1.0 : (ny == 2 ? -1.0 : 0.0));
                        ci = pt.coeff * (1 - sigma) * (ny == 1 ? 1.0 : (ny == 3 ? -1.0 : 0.0));
                        if (cr == 0 && ci == 0) {
                            continue; // contributes exactly zero
                        }
                    }
                    bucket[pauli_bucket(gr.xl, zyl, hmax)].push_back(PauliTileTerm{pt.zy & ~smask, cr, ci, zyl, 0});
                }
                for (int k = 0; k < nb; k++) {
                    gr.tb[k] = static_cast<int>(ttab.size());
                    gr.used |= bucket[k].empty() ? 0U : 1U << k;
                    ttab.insert(ttab.end(), bucket[k].begin(), bucket[k].end());
                }
                gr.tb[nb] = static_cast<int>(ttab.size());
                gtab.push_back(gr);
            }
            gspan.emplace_back(g0, gtab.size() - g0);
            tspan.emplace_back(t_first, ttab.size() - t_first);
        }
        std::vector<PauliTerm> wtab;
        std::vector<size_t> wstart;
        for (uint64_t x : wide) {
            wstart.push_back(wtab.size());
            wtab.insert(wtab.end(), groups[x].begin(), groups[x].end());
        }
        const size_t gb = gtab.size() * sizeof(PauliTileGroup), tb = ttab.size() * sizeof(PauliTileTerm);
        const size_t tb_off = (gb + 255) & ~size_t{255}, wb_off = (tb_off + tb + 255) & ~size_t{255};
        std::vector<unsigned char> blob(wb_off + wtab.size() * sizeof(PauliTerm) + 8);
        std::memcpy(blob.data(), gtab.data(), gb);
        std::memcpy(blob.data() + tb_off, ttab.data(), tb);
        std::memcpy(blob.data() + wb_off, wtab.data(), wtab.size() * sizeof(PauliTerm));
        const size_t nlaunch = passes.size() + wide.size();
        const size_t part_bytes = (nlaunch * kReduceBlocks + 1) * sizeof(double2);
        unsigned char *dev = nullptr;
        double2 *part = nullptr;
        if (ext) {
            const size_t part_off = (blob.size() + 255) & ~size_t{255};
            if (part_off + part_bytes > ext_bytes) {
                throw InvalidArg("the observable needs " + std::to_string(part_off + part_bytes) +
                                 " bytes of scratch; the allocation-free path has " + std::to_string(ext_bytes));
            }
            HIP_OK(hipMemcpyAsync(ext, blob.data(), blob.size(), hipMemcpyHostToDevice, h->stream));
            dev = ext;
            part = reinterpret_cast<double2 *>(ext + part_off);
        } else {
            dev = static_cast<unsigned char *>(stage(h, blob.data(), blob.size()));
            part = reinterpret_cast<double2 *>(reduce_buf(h, part_bytes));
        }
        const auto *dgroups = reinterpret_cast<const PauliTileGroup *>(dev);
        const auto *dtterms = reinterpret_cast<const PauliTileTerm *>(dev + tb_off);
        const auto *dwterms = reinterpret_cast<const PauliTerm *>(dev + wb_off);
        int total = 0;
        for (size_t pi = 0; pi < passes.size(); pi++) {
            const PauliTileGeom &g = geos[pi];
            const int nS = g.nlow + g.nhigh;
            const uint64_t ntiles = len() >> nS;
            const size_t lds = (size_t{1} << nS) * sizeof(V);
            const int ng = static_cast<int>(gspan[pi].second);
            const bool full = nS == 3 + hmax; // every state of >= 3 + hmax qubits
            auto launch = [&](auto kernel, int threads) {
                // persistent workgroups: as many as the device holds at once
                int per_cu = 0;
                HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
                const uint64_t resident = static_cast<uint64_t>(device_cus()) * std::max(per_cu, 1);
                const unsigned grid = static_cast<unsigned>(
                    std::min<uint64_t>({ntiles, resident, static_cast<uint64_t>(kReduceBlocks)}));
                hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), lds, h->stream, sv, ntiles, g, ng,
                                   dgroups + gspan[pi].first, dtterms, static_cast<int>(tspan[pi].first),
                                   static_cast<int>(tspan[pi].second), part + total);
                return grid;
            };
            // a pass of at most 64 terms takes its out-of-tile signs from one ballot per tile
            const bool signs = tspan[pi].second <= 64;
            auto pick = [&](auto high) -> unsigned {
                constexpr int H = decltype(high)::value;
                if (full) {
                    return signs ? launch(k_pauli_tile<T, H, true, true>, 1 << H)
                                 : launch(k_pauli_tile<T, H, true, false>, 1 << H);
                }
                return signs ? launch(k_pauli_tile<T, H, false, true>, 1 << H)
                             : launch(k_pauli_tile<T, H, false, false>, 1 << H);
            };
            const unsigned grid = hmax == kPauliTileHighMax ? pick(std::integral_constant<int, kPauliTileHighMax>{})
                                                            : pick(std::integral_constant<int, kPauliTileHigh>{});
            HIP_OK(hipGetLastError());
            total += grid;
        }
        for (size_t wi = 0; wi < wide.size(); wi++) {
            const uint64_t x = wide[wi];
            const unsigned grid = std::min<unsigned>(grid_for(len() / 2, 1, 4), kReduceBlocks);
            hipLaunchKernelGGL((k_pauli_expval<T>), dim3(grid), dim3(kBlock), 0, h->stream, sv, len() / 2, x,
                               __builtin_ctzll(x), static_cast<int>(groups[x].size()), dwterms + wstart[wi],
                               part + total);
            HIP_OK(hipGetLastError());
            total += grid;
        }
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, h->stream, part, total,
                           part + nlaunch * kReduceBlocks);
        HIP_OK(hipGetLastError());
        double2 r;
        HIP_OK(hipMemcpyAsync(&r, part + nlaunch * kReduceBlocks, sizeof(r), hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        return r.x;
    }

    // dst = P src (accumulate: dst += c P src)
    void apply_pauli(const V *src, V *dst, const PauliWord &pw, std::complex<double> c, bool acc) {
        // (-i)^ny
        static const std::complex<double> ph[4] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}};
        const std::complex<double> f = c * ph[pw.ny & 3];
        hipLaunchKernelGGL((k_apply_pauli<T>), dim3(grid_for(len())), dim3(kBlock), 0, h->stream,
                           src, dst, len(), pw.x, pw.zy,
                           dcx<T>{static_cast<T>(f.real()), static_cast<T>(f.imag())},
                           acc ? 1 : 0);
        HIP_OK(hipGetLastError());
    }

    void copy(V *dst, const V *src) {
        HIP_OK(hipMemcpyAsync(dst, src, h->state_bytes(), hipMemcpyDeviceToDevice, h->stream));
    }

    // Observable::applyInPlace (Observables.hpp:98-324, ObservablesLQubit.hpp:156-355)
    void apply_obs(V *sv, const ObsNode &o) {
        switch (o.kind) {
        case 0:
            gate(sv, o.op, o.wires.data(), static_cast<int>(o.wires.size()), false, nullptr, 0);
            return;
        case 1: {
            std::vector<double> m(o.mat.size() * 2);
            for (size_t i = 0; i < o.mat.size(); i++) {
                m[2 * i] = o.mat[i].real();
                m[2 * i + 1] = o.mat[i].imag();
            }
            matrix(sv, m.data(), o.wires.data(), static_cast<int>(o.wires.size()), false);
            return;
        }
        case 2:
            for (const auto &k : o.kids) {
                apply_obs(sv, k);
            }
            return;
        default: {
            // Hamiltonian: res = sum_t c_t (O_t psi), then psi <- res
            ScratchScope sc(h);
            V *res = static_cast<V *>(sc.get());
            bool first = true;
            V *tmp = nullptr;
            for (size_t t = 0; t < o.kids.size(); t++) {
                PauliWord pw;
                if (as_pauli(o.kids[t], n(), pw)) {
                    apply_pauli(sv, res, pw, {o.coeffs[t], 0}, !first);
                } else {
                    if (!tmp) {
                        tmp = static_cast<V *>(sc.get());
                    }
                    copy(tmp, sv);
                    apply_obs(tmp, o.kids[t]);
                    if (first) {
                        HIP_OK(hipMemsetAsync(res, 0, h->state_bytes(), h->stream));
                    }
                    hipLaunchKernelGGL((k_axpy<T>), dim3(grid_for(len())), dim3(kBlock), 0,
                                       h->stream, tmp, res, len(),
                                       dcx<T>{static_cast<T>(o.coeffs[t]), T(0)});
                    HIP_OK(hipGetLastError());
                }
                first = false;
            }
            if (first) {
                HIP_OK(hipMemsetAsync(res, 0, h->state_bytes(), h->stream));
            }
            copy(sv, res);
        }
        }
    }

    // out-of-place O psi into dst
    void obs_into(V *dst, const V *src, const ObsNode &o) {
        PauliWord pw;
        if (as_pauli(o, n(), pw)) {
            apply_pauli(src, dst, pw, {1, 0}, false);
            return;
        }
        if (o.kind == 3) {
            bool all = true;
            std::vector<PauliWord> ws(o.kids.size());
            for (size_t t = 0; t < o.kids.size() && all; t++) {
                all = as_pauli(o.kids[t], n(), ws[t]);
            }
            if (all && !o.kids.empty()) {
                for (size_t t = 0; t < ws.size(); t++) {
                    apply_pauli(src, dst, ws[t], {o.coeffs[t], 0}, t > 0);
                }
                return;
            }
        }
        copy(dst, src);
        apply_obs(dst, o);
    }

    // Measurements::probs() / probs(wires) (MeasurementsLQubit.hpp:79-135): the
    // marginal is computed directly in the reference's output axis order
    // (its sorted marginal followed by transpose_state_tensor).
    void probs(const int *wires, int nw, double *out) {
        if (nw < 0) {
            ScratchScope sc(h); // a state buffer holds >= 2^n doubles
            auto *d = static_cast<double *>(sc.get());
            hipLaunchKernelGGL((k_probs<T>), dim3(grid_for(len())), dim3(kBlock), 0, h->stream,
                               psi(), len(), d);
            HIP_OK(hipGetLastError());
            download_host(h, out, d, len() * sizeof(double));
            return;
        }
        if (nw > n() || nw > 32) {
            throw InvalidArg("too many wires for probs");
        }
        MarginalGeom m{};
        m.nw = nw;
        uint64_t pinned = 0;
        std::vector<int> bits;
        for (int j = 0; j < nw; j++) {
            if (wires[j] < 0 || wires[j] >= n()) {
                throw InvalidArg("wire index out of range");
            }
            const int b = n() - 1 - wires[j];
            if (pinned & (uint64_t{1} << b)) {
                throw InvalidArg("wires must be distinct");
            }
            pinned |= uint64_t{1} << b;
            bits.push_back(b);
        }
        std::sort(bits.begin(), bits.end());
        // Output axis j of the reference's result is wire sorted[argsort(wires)[j]]:
        // it marginalises over the sorted wires, then applies
        // transpose_state_tensor(probs, sorting_indices(wires))
        // (MeasurementsLQubit.hpp:99-135, Util.hpp:440-520).
        {
            std::vector<int> order(nw), sorted(wires, wires + nw);
            std::iota(order.begin(), order.end(), 0);
            std::sort(order.begin(), order.end(), [&](int a, int b) { return wires[a] < wires[b]; });
            std::sort(sorted.begin(), sorted.end());
            for (int j = 0; j < nw; j++) {
                m.bitpos[j] = n() - 1 - sorted[order[j]];
            }
        }
        m.free_mask = (len() - 1) & ~pinned;
        const uint64_t nbins = uint64_t{1} << nw, nrest = len() >> nw;
        double *d = nullptr;
        HIP_OK(hipMalloc(reinterpret_cast<void **>(&d), nbins * sizeof(double)));
        struct FreeD {
            double *p;
            ~FreeD() { (void)hipFree(p); }
        } free_d{d};
        if (nw <= 12 && nrest >= 4096) {
            m.rest.nb = nw;
            m.rest.fixed = 0;
            for (int j = 0; j < nw; j++) {
                m.rest.pos[j] = bits[j];
            }
            // marginal bits inside a 128-byte line: one block per pattern of
            // the others, every thread reading whole lines (k_marginal_lo)
            const uint64_t lomask = pinned & 7U, himask = pinned & ~uint64_t{7};
            const int lo = __builtin_popcountll(lomask);
            const uint64_t nys = nbins >> lo; // blocks along y
            const unsigned chunks = static_cast<unsigned>(std::max<uint64_t>(
                1, std::min<uint64_t>((nrest + kBlock * 64 - 1) / (kBlock * 64),
                                      std::max<uint64_t>(1, uint64_t(device_cus()) * 16 / nys))));
            double *part = reduce_buf(h, nbins * chunks * sizeof(double));
            const dim3 grid(chunks, static_cast<unsigned>(nys));
            switch (lo) {
            case 0: hipLaunchKernelGGL((k_marginal<T>), grid, dim3(kBlock), 0, h->stream, psi(), nrest, m, part); break;
            case 1: hipLaunchKernelGGL((k_marginal_lo<T, 1>), grid, dim3(kBlock), 0, h->stream, psi(), nrest, m, lomask, himask, part); break;
            case 2: hipLaunchKernelGGL((k_marginal_lo<T, 2>), grid, dim3(kBlock), 0, h->stream, psi(), nrest, m, lomask, himask, part); break;
            default: hipLaunchKernelGGL((k_marginal_lo<T, 3>), grid, dim3(kBlock), 0, h->stream, psi(), nrest, m, lomask, himask, part); break;
            }
            HIP_OK(hipGetLastError());
            hipLaunchKernelGGL(k_marginal_final, dim3((nbins + kBlock - 1) / kBlock), dim3(kBlock), 0,
                               h->stream, part, nbins, (int)chunks, d);
        } else {
            hipLaunchKernelGGL((k_marginal_small<T>), dim3((nbins + kBlock - 1) / kBlock),
                               dim3(kBlock), 0, h->stream, psi(), nbins, nrest, m, d);
        }
        HIP_OK(hipGetLastError());
        download_host(h, out, d, nbins * sizeof(double));
    }

    // Exclusive-prefix tables of |psi|^2 for inverse-CDF sampling
    // (kernels_sample.hpp), in the handle's buffers, for the current state.
    struct CdfTables {
        const double *gpre, *tpre;
        uint64_t ngroups, ntiles;
    };
    CdfTables cdf_tables() {
        const uint64_t ngroups = (len() + kCdfGroup - 1) / kCdfGroup;
        const uint64_t ntiles = (ngroups + kCdfTile - 1) / kCdfTile;
        auto *gsum = static_cast<double *>(buf(h, kBufCdfGroups, ngroups * sizeof(double)));
        auto *gpre = static_cast<double *>(buf(h, kBufCdfPrefix, ngroups * sizeof(double)));
        auto *tsum = static_cast<double *>(buf(h, kBufCdfTiles, (2 * ntiles + 1) * sizeof(double)));
        double *tpre = tsum + ntiles; // ntiles + 1 entries, the last one the total
        const uint64_t nwaves = (ngroups + 4 * kCdfSteps - 1) / (4 * kCdfSteps);
        hipLaunchKernelGGL((k_cdf_groups<T>), dim3(static_cast<unsigned>((nwaves + 3) / 4)), dim3(kBlock), 0,
                           h->stream, psi(), len(), ngroups, gsum);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(k_cdf_tile_sums, dim3(static_cast<unsigned>(ntiles)), dim3(kBlock), 0, h->stream,
                           gsum, ngroups, tsum);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, h->stream, tsum, ntiles, tpre);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(k_cdf_tile_scan, dim3(static_cast<unsigned>(ntiles)), dim3(kBlock), 0, h->stream,
                           gsum, ngroups, tpre, gpre);
        HIP_OK(hipGetLastError());
        return {gpre, tpre, ngroups, ntiles};
    }
    // didx[s] for the uniforms du[s], both in device memory (queued, no wait)
    void sample_device(const double *du, int64_t m, int64_t *didx) {
        const CdfTables c = cdf_tables();
        hipLaunchKernelGGL((k_cdf_search<T>), dim3(static_cast<unsigned>((m + kBlock - 1) / kBlock)), dim3(kBlock),
                           0, h->stream, psi(), len(), c.ngroups, c.ntiles, c.tpre, c.gpre, du, m, didx);
        HIP_OK(hipGetLastError());
    }

    // Inverse-CDF sampling.  Either host uniforms (-> indices) or a
    // SplitMix64 seed (-> bit rows).
    void sample(const double *host_u, uint64_t seed, int64_t m, int64_t *host_idx,
                uint64_t *host_bits) {
        auto *du = static_cast<double *>(buf(h, kBufUniforms, m * sizeof(double)));
        auto *didx = static_cast<int64_t *>(buf(h, kBufIndices, m * sizeof(int64_t)));
        if (host_u) {
            upload_host(h, du, host_u, m * sizeof(double));
        } else {
            hipLaunchKernelGGL(k_uniforms, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0,
                               h->stream, seed, m, du);
            HIP_OK(hipGetLastError());
        }
        sample_device(du, m, didx);
        if (host_idx) {
            download_host(h, host_idx, didx, m * sizeof(int64_t));
        }
        if (host_bits) {
            const uint64_t tot = uint64_t(m) * n();
            auto *dbits = static_cast<uint64_t *>(buf(h, kBufBits, tot * sizeof(uint64_t)));
            hipLaunchKernelGGL(k_sample_bits, dim3((tot + kBlock - 1) / kBlock), dim3(kBlock), 0,
                               h->stream, didx, m, n(), dbits);
            HIP_OK(hipGetLastError());
            download_host(h, host_bits, dbits, tot * sizeof(uint64_t));
        }
        HIP_OK(hipStreamSynchronize(h->stream));
    }

    // Metropolis-Hastings samples (kernels_mcmc.hpp); bits as generate_samples
    void mcmc(int kind, int64_t burnin, int64_t m, uint64_t seed, int nchains, uint64_t *host_bits) {
        if (kind != 0 && kind != 1) {
            throw InvalidArg("transition kernel must be 0 (non-zero uniform) or 1 (local flip)");
        }
        if (burnin < 0 || m < 1) {
            throw InvalidArg("need num_burnin >= 0 and num_shots >= 1");
        }
        struct Free {
            std::vector<void *> p;
            ~Free() {
                for (void *q : p) {
                    (void)hipFree(q);
                }
            }
        } fr;
        auto alloc = [&](size_t bytes) {
            void *q = nullptr;
            HIP_OK(hipMalloc(&q, std::max<size_t>(bytes, 8)));
            fr.p.push_back(q);
            return q;
        };
        uint64_t *nz = nullptr;
        uint64_t nnz = len();
        if (kind == 0) {
            const uint64_t nb = (len() + kNzChunk - 1) / kNzChunk;
            auto *counts = static_cast<double *>(alloc(nb * sizeof(double)));
            auto *offs = static_cast<double *>(alloc((nb + 1) * sizeof(double)));
            const T eps = std::numeric_limits<T>::epsilon();
            hipLaunchKernelGGL((k_nz_count<T>), dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, h->stream,
                               psi(), len(), eps, counts);
            HIP_OK(hipGetLastError());
            hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, h->stream, counts, nb, offs);
            HIP_OK(hipGetLastError());
            double total = 0;
            HIP_OK(hipMemcpyAsync(&total, offs + nb, sizeof(double), hipMemcpyDeviceToHost, h->stream));
            HIP_OK(hipStreamSynchronize(h->stream));
            nnz = static_cast<uint64_t>(total);
            if (nnz == 0) {
                throw InvalidArg("the state has no amplitude above the precision's epsilon");
            }
            if (nnz < len()) {
                nz = static_cast<uint64_t *>(alloc(nnz * sizeof(uint64_t)));
                hipLaunchKernelGGL((k_nz_write<T>), dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, h->stream,
                                   psi(), len(), eps, offs, nz);
                HIP_OK(hipGetLastError());
            }
        }
        const uint64_t chains = nchains > 0 ? std::min<uint64_t>(nchains, m)
                                            : std::min<uint64_t>(m, uint64_t{16384});
        const uint64_t per = (m + chains - 1) / chains;
        // every chain starts from an inverse-CDF draw, i.e. in the stationary
        // distribution |psi|^2, so that short chains are still unbiased
        auto *ustart = static_cast<double *>(buf(h, kBufUniforms, chains * sizeof(double)));
        auto *start = static_cast<int64_t *>(buf(h, kBufChainStart, chains * sizeof(int64_t)));
        hipLaunchKernelGGL(k_uniforms, dim3(static_cast<unsigned>((chains + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           h->stream, seed ^ 0xA0761D6478BD642Full, static_cast<int64_t>(chains), ustart);
        HIP_OK(hipGetLastError());
        sample_device(ustart, static_cast<int64_t>(chains), start);
        auto *didx = static_cast<int64_t *>(alloc(m * sizeof(int64_t)));
        hipLaunchKernelGGL((k_mcmc_chains<T>), dim3(static_cast<unsigned>((chains + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, h->stream, psi(), len(), n(), kind, nz, nnz, start,
                           static_cast<uint64_t>(burnin), per, static_cast<uint64_t>(m), seed, didx);
        HIP_OK(hipGetLastError());
        const uint64_t tot = uint64_t(m) * n();
        auto *dbits = static_cast<uint64_t *>(alloc(tot * sizeof(uint64_t)));
        hipLaunchKernelGGL(k_sample_bits, dim3((tot + kBlock - 1) / kBlock), dim3(kBlock), 0, h->stream, didx, m,
                           n(), dbits);
        HIP_OK(hipGetLastError());
        download_host(h, host_bits, dbits, tot * sizeof(uint64_t));
    }

    // Measurements::expval(Observable) and var(Observable)
    // a Pauli word or a Hamiltonian of Pauli words as (words, coefficients)
    bool as_pauli_sum(const ObsNode &o, std::vector<PauliWord> &ws, std::vector<double> &cs) {
        PauliWord pw;
        if (as_pauli(o, n(), pw)) {
            ws = {pw};
            cs = {1.0};
            return true;
        }
        if (o.kind != 3 || o.kids.empty()) {
            return false;
        }
        ws.assign(o.kids.size(), PauliWord{});
        for (size_t t = 0; t < o.kids.size(); t++) {
            if (!as_pauli(o.kids[t], n(), ws[t])) {
                return false;
            }
        }
        cs = o.coeffs;
        return true;
    }

    double expval(const ObsNode &o, bool var) {
        if (!var) {
            std::vector<PauliWord> ws;
            std::vector<double> cs;
            if (as_pauli_sum(o, ws, cs)) {
                return pauli_expval(psi(), ws, cs);
            }
        }
        ScratchScope sc(h);
        V *bra = static_cast<V *>(sc.get());
        obs_into(bra, psi(), o);
        if (!var) {
            return dot(bra, psi()).real();
        }
        const double ms = dot(bra, bra).real();
        const double e = dot(bra, psi()).real();
        return ms - e * e;
    }
};

// ---- deferred gates --------------------------------------------------------
// While deferral is on, pl_sv_apply_gate / pl_sv_apply_ops check each gate as
// the per-gate path would (same errors, at the call) and queue it; the queue
// is applied as tile passes (the precision's default tiles) by the first call
// that reads or changes the state otherwise, or once it holds
// kDeferMaxOps gates.
constexpr size_t kDeferMaxOps = 4096;

void flush_deferred(pl_sv_t h) {
    if (!h || h->q_ids.empty()) {
        return;
    }
    // the queue leaves the handle first: a failure cannot apply it twice
    std::vector<int> ids, nw, wires, np;
    std::vector<uint8_t> inv;
    std::vector<double> prm;
    ids.swap(h->q_ids);
    nw.swap(h->q_nwires);
    wires.swap(h->q_wires);
    np.swap(h->q_nparams);
    inv.swap(h->q_inverse);
    prm.swap(h->q_params);
    const TapeView t{static_cast<int>(ids.size()), ids.data(), nw.data(), wires.data(), inv.data(), np.data(),
                     prm.data()};
    HIP_OK(hipSetDevice(0));
    if (h->dtype == PL_C128) {
        Engine<double> e(h);
        e.apply_tiled(e.psi(), t, 0, -1);
    } else {
        Engine<float> e(h);
        e.apply_tiled(e.psi(), t, 0, -1);
    }
}

// the queue dropped: the state is about to be overwritten whole
void discard_deferred(pl_sv_t h) {
    h->q_ids.clear();
    h->q_nwires.clear();
    h->q_wires.clear();
    h->q_nparams.clear();
    h->q_inverse.clear();
    h->q_params.clear();
}

void defer_gate(pl_sv_t h, int id, const int *wires, int nw, bool inverse, const double *params, int np) {
    (void)lower_gate<double>(h->n, id, wires, nw, inverse, params, np); // the per-gate path's checks
    h->q_ids.push_back(id);
    h->q_nwires.push_back(nw);
    h->q_wires.insert(h->q_wires.end(), wires, wires + nw);
    h->q_inverse.push_back(inverse ? 1 : 0);
    h->q_nparams.push_back(np);
    h->q_params.insert(h->q_params.end(), params, params + np);
    if (h->q_ids.size() >= kDeferMaxOps) {
        flush_deferred(h);
    }
}

// Engine of the handle's precision, exceptions propagate
template <class F> void dispatch_or_throw(pl_sv_t h, F &&f) {
    flush_deferred(h);
    HIP_OK(hipSetDevice(0));
    if (h->dtype == PL_C128) {
        Engine<double> e(h);
        f(e);
    } else {
        Engine<float> e(h);
        f(e);
    }
}

template <class F> int dispatch(pl_sv_t h, F &&f) {
    return guarded([&] {
        if (!h) {
            throw InvalidArg("null state-vector handle");
        }
        flush_deferred(h);
        HIP_OK(hipSetDevice(0));
        if (h->dtype == PL_C128) {
            Engine<double> e(h);
            f(e);
        } else {
            Engine<float> e(h);
            f(e);
        }
    });
}

// ---------------------------------------------------------------------------
// adjoint sweep (AdjointJacobianLQubit.hpp:212-336)
// ---------------------------------------------------------------------------
template <class T>
void adjoint_sweep(Engine<T> &E, int nops, const int *ids, const int *nwires, const int *wires,
                   const uint8_t *inverse, const int *nparams, const double *params,
                   const int *prog, const double *coeffs, const double *mats, int ntp,
                   const int *tp, double *jac_out) {
    using V = v2<T>;
    pl_sv *h = E.h;
    struct Op {
        int id;
        std::vector<int> w;
        bool inv;
        std::vector<double> p;
    };
    std::vector<Op> ops(nops);
    for (int i = 0, wo = 0, po = 0; i < nops; i++) {
        ops[i].id = ids[i];
        ops[i].w.assign(wires + wo, wires + wo + nwires[i]);
        wo += nwires[i];
        ops[i].inv = inverse[i] != 0;
        ops[i].p.assign(params + po, params + po + nparams[i]);
        po += nparams[i];
        if (ops[i].id != -1) {
            gate_name(ops[i].id);
        }
    }
    ObsReader rd{prog, coeffs, mats};
    const int K = *rd.prog++;
    std::vector<ObsNode> obs;
    for (int k = 0; k < K; k++) {
        obs.push_back(rd.one());
    }
    if (ntp == 0) {
        return; // JacobianData::hasTrainableParams() false: nothing to do
    }
    for (int t = 1; t < ntp; t++) {
        if (tp[t] <= tp[t - 1]) {
            throw InvalidArg("trainable parameter indices must be sorted ascending");
        }
    }
    if (K == 0) {
        return;
    }
    ScratchScope sc(h);
    V *lam = static_cast<V *>(sc.get());
    E.copy(lam, E.psi());
    std::vector<V *> states{lam};
    for (int k = 0; k < K; k++) {
        V *hk = static_cast<V *>(sc.get());
        E.obs_into(hk, lam, obs[k]);
        states.push_back(hk);
    }
    // Observables per sweep: when lambda and every H_k exceed the on-die
    // cache but fit in it in groups, the sweep runs once per group with a
    // fresh lambda, so each group's states stay cache-resident across ops.
    const size_t sbytes = E.len() * sizeof(V);
    int per_sweep = K;
    if (sbytes * (K + 1) > kAdjCacheBytes && sbytes * 2 <= kAdjCacheBytes) {
        const int fit = std::max<int>(1, static_cast<int>(kAdjCacheBytes / sbytes) - 1);
        const int nsweep = (K + fit - 1) / fit;
        per_sweep = (K + nsweep - 1) / nsweep;
    }
    // device tables: state pointers, jacobian, partials
    V **dstates = nullptr;
    double *djac = nullptr, *dpart = nullptr;
    const unsigned grid = std::min<unsigned>(grid_for(E.len(), 1, 8), kReduceBlocks);
    // one pointer table per chunk of at most kAdjMaxObs observables of a
    // sweep: lambda, then the chunk's H_k (built once, so no launch waits on
    // a staging copy); chunks are numbered sweep by sweep
    std::vector<V *> tables;
    for (int s0 = 0; s0 < K; s0 += per_sweep) {
        const int ns = std::min(per_sweep, K - s0);
        for (int k0 = 0; k0 < ns; k0 += kAdjMaxObs) {
            const size_t at = tables.size();
            tables.resize(at + kAdjMaxObs + 1, nullptr);
            tables[at] = lam;
            for (int k = k0; k < std::min(ns, k0 + kAdjMaxObs); k++) {
                tables[at + 1 + (k - k0)] = states[1 + s0 + k];
            }
        }
    }
    HIP_OK(hipMalloc(&dstates, tables.size() * sizeof(V *)));
    HIP_OK(hipMalloc(&djac, size_t(ntp) * K * sizeof(double)));
    // each trainable step's partial sums in a slot of their own; the slots'
    // final sums run kAdjFinalSlots at a time (one launch, not one per step)
    const size_t slot_stride = size_t(grid) * kAdjMaxObs;
    HIP_OK(hipMalloc(&dpart, kAdjFinalSlots * slot_stride * sizeof(double)));
    struct Free {
        void *a, *b, *c;
        ~Free() {
            (void)hipFree(a);
            (void)hipFree(b);
            (void)hipFree(c);
        }
    } fr{dstates, djac, dpart};
    HIP_OK(hipMemcpyAsync(dstates, tables.data(), tables.size() * sizeof(V *),
                          hipMemcpyHostToDevice, E.h->stream));
    HIP_OK(hipMemsetAsync(djac, 0, size_t(ntp) * K * sizeof(double), E.h->stream));

    size_t num_par_ops = 0;
    for (auto &o : ops) {
        num_par_ops += o.p.empty() ? 0 : 1;
    }
    AdjFinals pending{};
    int npending = 0;
    auto flush_finals = [&] {
        if (npending > 0) {
            hipLaunchKernelGGL(k_adjoint_final_many, dim3(kAdjMaxObs, npending), dim3(kBlock), 0, E.h->stream,
                               dpart, slot_stride, pending, djac);
            HIP_OK(hipGetLastError());
            npending = 0;
        }
    };
    // the slot for the next trainable step's partials (all slots in use: their
    // sums first)
    auto slot = [&] {
        if (npending == kAdjFinalSlots) {
            flush_finals();
        }
        return dpart + static_cast<size_t>(npending) * slot_stride;
    };
    auto finish = [&](unsigned nblocks, int nk, double factor, int dst) {
        pending.t[npending++] = AdjFinal{factor, static_cast<int>(nblocks), nk, dst};
    };
    ScratchScope sc2(h);
    V *mu = nullptr;
    size_t chunk_base = 0; // first pointer table of the current sweep
    for (int s0 = 0; s0 < K; s0 += per_sweep) {
    const int ns = std::min(per_sweep, K - s0);
    if (s0 > 0) {
        E.copy(lam, E.psi()); // the next group starts again from psi
    }
    int t_num = ntp - 1;
    long long cur_param = static_cast<long long>(num_par_ops) - 1;
    int tp_pos = ntp - 1;
    for (int i = nops - 1; i >= 0; i--) {
        const Op &op = ops[i];
        if (op.p.size() > 1) {
            throw InvalidArg("The operation is not supported using the adjoint differentiation method");
        }
        if (op.id == -1) {
            continue;
        }
        if (tp_pos < 0) {
            break;
        }
        const auto pt = E.params_t(op.p.data(), static_cast<int>(op.p.size()));
        const GroupOp<T> ud = lower_gate<T>(E.n(), op.id, op.w.data(), static_cast<int>(op.w.size()),
                                            !op.inv, pt.data(), static_cast<int>(pt.size()));
        bool train = false;
        GroupOp<T> g;
        double factor = 0;
        int row = 0;
        if (!op.p.empty()) {
            if (cur_param == tp[tp_pos]) {
                train = true;
                T s = 0;
                g = lower_generator<T>(E.n(), op.id, op.w.data(), static_cast<int>(op.w.size()),
                                       true, &s);
                factor = -2.0 * static_cast<double>(s) * (op.inv ? -1.0 : 1.0);
                row = t_num * K + s0;
                t_num--;
                tp_pos--;
            }
            cur_param--;
        }
        const bool fusable =
            ud.kind == GroupOp<T>::Noop ||
            (ud.kind == GroupOp<T>::Parity && (!train || g.same_geometry(ud))) ||
            (ud.kind == GroupOp<T>::Group && ud.M <= 4 && (!train || g.same_geometry(ud)));
        if (!fusable) {
            // unfused: mu = lambda; lambda <- U^dag lambda; mu <- G mu; dots; H_k <- U^dag H_k
            if (train) {
                if (!mu) {
                    mu = static_cast<V *>(sc2.get());
                }
                E.copy(mu, lam);
            }
            E.apply(ud, lam);
            if (train) {
                T s = 0;
                E.apply(lower_generator<T>(E.n(), op.id, op.w.data(), static_cast<int>(op.w.size()),
                                           false, &s),
                        mu);
                std::vector<double> row_vals(ns);
                for (int k = 0; k < ns; k++) {
                    row_vals[k] = factor * E.dot(states[1 + s0 + k], mu).imag();
                }
                HIP_OK(hipMemcpyAsync(djac + row, row_vals.data(), ns * sizeof(double),
                                      hipMemcpyHostToDevice, E.h->stream));
                HIP_OK(hipStreamSynchronize(E.h->stream));
            }
            for (int k = 0; k < ns; k++) {
                E.apply(ud, states[1 + s0 + k]);
            }
            continue;
        }
        if (ud.kind == GroupOp<T>::Noop) {
            // Identity: U^dag = I, no generator
            continue;
        }
        // fused launches, kAdjMaxObs observables each; lambda updated by the last
        for (int k0 = 0; k0 < ns; k0 += kAdjMaxObs) {
            const int nk = std::min(kAdjMaxObs, ns - k0);
            const int upd = (k0 + nk == ns) ? 1 : 0;
            // states for this launch: lambda followed by H_(s0+k0) .. H_(s0+k0+nk-1)
            V *const *st = dstates + (chunk_base + size_t(k0 / kAdjMaxObs)) * (kAdjMaxObs + 1);
            if (ud.kind == GroupOp<T>::Parity) {
                const dcx<T> u0{ud.parity_phase[0].real(), ud.parity_phase[0].imag()};
                const dcx<T> u1{ud.parity_phase[1].real(), ud.parity_phase[1].imag()};
                dcx<T> g0{1, 0}, g1{1, 0};
                if (train) {
                    g0 = {g.parity_phase[0].real(), g.parity_phase[0].imag()};
                    g1 = {g.parity_phase[1].real(), g.parity_phase[1].imag()};
                    hipLaunchKernelGGL((k_adjoint_parity<T, true>), dim3(grid), dim3(kBlock), 0,
                                       E.h->stream, st, nk, E.len(), ud.parity_mask, u0, u1, g0,
                                       g1, upd, slot());
                } else {
                    hipLaunchKernelGGL((k_adjoint_parity<T, false>), dim3(grid), dim3(kBlock), 0,
                                       E.h->stream, st, nk, E.len(), ud.parity_mask, u0, u1, g0,
                                       g1, upd, dpart);
                }
            } else {
                auto go = [&](auto mtag) {
                    constexpr int M = decltype(mtag)::value;
                    AdjOp<T, M> a{};
                    a.geo.fixed = ud.fixed;
                    a.geo.nb = ud.nb;
                    for (int q = 0; q < ud.nb; q++) {
                        a.geo.pos[q] = ud.pos[q];
                    }
                    for (int j = 0; j < M; j++) {
                        a.off[j] = ud.off[j];
                    }
                    for (int j = 0; j < M * M; j++) {
                        a.ud[j] = {ud.mat[j].real(), ud.mat[j].imag()};
                        a.g[j] = train ? dcx<T>{g.mat[j].real(), g.mat[j].imag()} : dcx<T>{0, 0};
                    }
                    a.ud_diag = ud.diag ? 1 : 0;
                    const uint64_t nwork = E.len() >> ud.nb;
                    // a trainable step ends with one cross-lane reduction per observable
                    // in every wavefront: two work items per thread amortise it
                    const unsigned gr = std::min<unsigned>(grid_for(nwork, train ? 2 : 1, 8), kReduceBlocks);
                    if (train) {
                        hipLaunchKernelGGL((k_adjoint_step<T, M, true>), dim3(gr), dim3(kBlock), 0,
                                           E.h->stream, st, nk, nwork, a, upd, slot());
                    } else {
                        hipLaunchKernelGGL((k_adjoint_step<T, M, false>), dim3(gr), dim3(kBlock), 0,
                                           E.h->stream, st, nk, nwork, a, upd, dpart);
                    }
                    return gr;
                };
                unsigned gr = 0;
                switch (ud.M) {
                case 1: gr = go(std::integral_constant<int, 1>{}); break;
                case 2: gr = go(std::integral_constant<int, 2>{}); break;
                default: gr = go(std::integral_constant<int, 4>{}); break;
                }
                HIP_OK(hipGetLastError());
                if (train) {
                    finish(gr, nk, factor, row + k0);
                }
                continue;
            }
            HIP_OK(hipGetLastError());
            if (train) {
                finish(grid, nk, factor, row + k0);
            }
        }
    }
    chunk_base += static_cast<size_t>((ns + kAdjMaxObs - 1) / kAdjMaxObs);
    }
    flush_finals();
    std::vector<double> J(size_t(ntp) * K);
    HIP_OK(hipMemcpyAsync(J.data(), djac, J.size() * sizeof(double), hipMemcpyDeviceToHost,
                          E.h->stream));
    HIP_OK(hipStreamSynchronize(E.h->stream));
    // parameter-major -> observable-major (AdjointJacobianLQubit.hpp:331-334)
    for (int k = 0; k < K; k++) {
        for (int t = 0; t < ntp; t++) {
            jac_out[size_t(k) * ntp + t] = J[size_t(t) * K + k];
        }
    }
}

} // namespace plamd

using namespace plamd;

// ===========================================================================
// extern "C"
// ===========================================================================
// VectorJacobianProduct::operator() (VectorJacobianProduct.hpp:77-155):
// lambda = psi, mu = dy; for operations in reverse (state preparations
// skipped, stopping once every trainable parameter is done): for a trainable
// one, w_t = i s <G mu | lambda> with s the generator scale times -1 for an
// inverted operation; then lambda <- U^dag lambda and mu <- U^dag mu.  Ops of
// at most 4 members (and parity phases) whose generator shares their
// geometry take one fused launch each (k_vjp_step); the rest go op by op.
template <class T>
void vjp_sweep(Engine<T> &E, int nops, const int *ids, const int *nwires, const int *wires,
               const uint8_t *inverse, const int *nparams, const double *params, const double *dy,
               int ntp, const int *tp, double *out) {
    using V = v2<T>;
    pl_sv *h = E.h;
    std::vector<int> woff(nops + 1, 0), poff(nops + 1, 0);
    size_t num_par_ops = 0;
    for (int i = 0; i < nops; i++) {
        woff[i + 1] = woff[i] + nwires[i];
        poff[i + 1] = poff[i] + nparams[i];
        num_par_ops += nparams[i] > 0 ? 1 : 0;
        if (ids[i] != -1) {
            gate_name(ids[i]);
        }
    }
    for (int t = 1; t < ntp; t++) {
        if (tp[t] <= tp[t - 1]) {
            throw InvalidArg("trainable parameter indices must be sorted ascending");
        }
    }
    if (ntp == 0) {
        return;
    }
    ScratchScope sc(h);
    V *lam = static_cast<V *>(sc.get());
    V *mu = static_cast<V *>(sc.get());
    V *mud = static_cast<V *>(sc.get());
    E.copy(lam, E.psi());
    {
        std::vector<V> host(E.len());
        for (uint64_t i = 0; i < E.len(); i++) {
            host[i].x = static_cast<T>(dy[2 * i]);
            host[i].y = static_cast<T>(dy[2 * i + 1]);
        }
        upload_host(h, mu, host.data(), E.len() * sizeof(V));
        HIP_OK(hipStreamSynchronize(h->stream));
    }
    long long cur_param = static_cast<long long>(num_par_ops) - 1;
    int tp_pos = ntp - 1;
    // <G mu|lambda> of trainable parameter t stays on the device in dots[t]
    // until the sweep ends: one read-back instead of a wait per parameter
    auto *part = reinterpret_cast<double2 *>(reduce_buf(h, (kReduceBlocks + size_t(ntp)) * sizeof(double2)));
    double2 *dots = part + kReduceBlocks;
    std::vector<double> scale(ntp, 0.0);
    std::vector<char> reached(ntp, 0);
    // fused steps (k_vjp_step / k_vjp_parity): one launch per op; a trainable
    // op's block partials go to its row of vpart, all rows reduced at the end
    const unsigned grid = std::min<unsigned>(grid_for(E.len(), 1, 8), kReduceBlocks);
    auto *vpart = static_cast<double2 *>(buf(h, kBufVjp, size_t(ntp) * grid * sizeof(double2)));
    std::vector<int> fused_rows;
    for (int i = nops - 1; i >= 0; i--) {
        if (nparams[i] > 1) {
            throw InvalidArg("The operation is not supported using the adjoint differentiation method");
        }
        if (ids[i] == -1) {
            continue;
        }
        if (tp_pos < 0) {
            break;
        }
        const int *w = wires + woff[i];
        const auto pt = E.params_t(params + poff[i], nparams[i]);
        const GroupOp<T> ud = lower_gate<T>(E.n(), ids[i], w, nwires[i], inverse[i] == 0, pt.data(),
                                            nparams[i]);
        bool train = false;
        int row = -1;
        GroupOp<T> g;
        if (nparams[i] > 0) {
            if (cur_param == tp[tp_pos]) {
                train = true;
                row = tp_pos;
                T s = 0;
                g = lower_generator<T>(E.n(), ids[i], w, nwires[i], true, &s);
                scale[tp_pos] = static_cast<double>(s) * (inverse[i] ? -1.0 : 1.0);
                reached[tp_pos] = 1;
                tp_pos--;
            }
            cur_param--;
        }
        const bool fusable = (ud.kind == GroupOp<T>::Noop && !train) ||
                             (ud.kind == GroupOp<T>::Parity && (!train || g.same_geometry(ud))) ||
                             (ud.kind == GroupOp<T>::Group && ud.M <= 4 && (!train || g.same_geometry(ud)));
        if (!fusable) {
            if (train) {
                E.copy(mud, mu);
                T s = 0;
                E.apply(lower_generator<T>(E.n(), ids[i], w, nwires[i], false, &s), mud);
                E.dot_to(mud, lam, part, dots + row);
            }
            E.apply(ud, lam);
            E.apply(ud, mu);
            continue;
        }
        if (ud.kind == GroupOp<T>::Noop) {
            continue;
        }
        double2 *pp = train ? vpart + size_t(row) * grid : nullptr;
        if (ud.kind == GroupOp<T>::Parity) {
            const dcx<T> u0{ud.parity_phase[0].real(), ud.parity_phase[0].imag()};
            const dcx<T> u1{ud.parity_phase[1].real(), ud.parity_phase[1].imag()};
            dcx<T> g0{1, 0}, g1{1, 0};
            if (train) {
                g0 = {g.parity_phase[0].real(), g.parity_phase[0].imag()};
                g1 = {g.parity_phase[1].real(), g.parity_phase[1].imag()};
                hipLaunchKernelGGL((k_vjp_parity<T, true>), dim3(grid), dim3(kBlock), 0, h->stream, lam, mu,
                                   E.len(), ud.parity_mask, u0, u1, g0, g1, pp);
            } else {
                hipLaunchKernelGGL((k_vjp_parity<T, false>), dim3(grid), dim3(kBlock), 0, h->stream, lam, mu,
                                   E.len(), ud.parity_mask, u0, u1, g0, g1, pp);
            }
        } else {
            auto go = [&](auto mtag) {
                constexpr int M = decltype(mtag)::value;
                AdjOp<T, M> a{};
                a.geo.fixed = ud.fixed;
                a.geo.nb = ud.nb;
                for (int q = 0; q < ud.nb; q++) {
                    a.geo.pos[q] = ud.pos[q];
                }
                for (int j = 0; j < M; j++) {
                    a.off[j] = ud.off[j];
                }
                for (int j = 0; j < M * M; j++) {
                    a.ud[j] = {ud.mat[j].real(), ud.mat[j].imag()};
                    a.g[j] = train ? dcx<T>{g.mat[j].real(), g.mat[j].imag()} : dcx<T>{0, 0};
                }
                a.ud_diag = ud.diag ? 1 : 0;
                const uint64_t nwork = E.len() >> ud.nb;
                if (train) {
                    hipLaunchKernelGGL((k_vjp_step<T, M, true>), dim3(grid), dim3(kBlock), 0, h->stream, lam, mu,
                                       nwork, a, pp);
                } else {
                    hipLaunchKernelGGL((k_vjp_step<T, M, false>), dim3(grid), dim3(kBlock), 0, h->stream, lam, mu,
                                       nwork, a, pp);
                }
            };
            switch (ud.M) {
            case 1: go(std::integral_constant<int, 1>{}); break;
            case 2: go(std::integral_constant<int, 2>{}); break;
            default: go(std::integral_constant<int, 4>{}); break;
            }
        }
        HIP_OK(hipGetLastError());
        if (train) {
            fused_rows.push_back(row);
        }
    }
    if (!fused_rows.empty()) {
        const auto *drows = static_cast<const int *>(stage(h, fused_rows.data(), fused_rows.size() * sizeof(int)));
        hipLaunchKernelGGL(k_vjp_final, dim3(static_cast<unsigned>(fused_rows.size())), dim3(kBlock), 0, h->stream,
                           vpart, static_cast<int>(grid), drows, dots);
        HIP_OK(hipGetLastError());
    }
    std::vector<double2> z(ntp);
    HIP_OK(hipMemcpyAsync(z.data(), dots, ntp * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    for (int t = 0; t < ntp; t++) {
        if (reached[t]) {
            out[2 * t] = -scale[t] * z[t].y;
            out[2 * t + 1] = scale[t] * z[t].x;
        }
    }
}

extern "C" {

const char *pl_last_error(void) { return g_err.c_str(); }

int pl_device_info(char *buf, size_t len) {
    return guarded([&] {
        int dev = 0;
        hipDeviceProp_t p;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipGetDeviceProperties(&p, dev));
        std::snprintf(buf, len, "{\"name\": \"%s\", \"arch\": \"%s\", \"cus\": %d, \"hbm_bytes\": %zu}",
                      p.name, p.gcnArchName, p.multiProcessorCount, p.totalGlobalMem);
    });
}

int pl_sv_create(int num_qubits, int dtype, pl_sv_t *out) {
    return guarded([&] {
        if (!out) {
            throw InvalidArg("null output handle");
        }
        if (num_qubits < 1 || num_qubits > 40) {
            throw InvalidArg("num_qubits must be in [1, 40]");
        }
        if (dtype != PL_C64 && dtype != PL_C128) {
            throw InvalidArg("dtype must be PL_C64 or PL_C128");
        }
        HIP_OK(hipSetDevice(0));
        auto h = std::make_unique<pl_sv>();
        h->n = num_qubits;
        h->dtype = dtype;
        h->len = uint64_t{1} << num_qubits;
        HIP_OK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        HIP_OK(hipEventCreate(&h->ev0));
        HIP_OK(hipEventCreate(&h->ev1));
        HIP_OK(hipEventCreateWithFlags(&h->ev_order, hipEventDisableTiming));
        hipError_t e = hipMalloc(&h->d, h->state_bytes());
        if (e != hipSuccess) {
            (void)hipStreamDestroy(h->stream);
            throw DeviceError(std::string("hipMalloc of the state failed: ") + hipGetErrorString(e));
        }
        pl_sv *raw = h.release();
        if (pl_sv_reset(raw) != PL_OK) {
            const std::string msg = g_err;
            pl_sv_destroy(raw);
            throw DeviceError(msg);
        }
        *out = raw;
    });
}

int pl_sv_destroy(pl_sv_t h) {
    return guarded([&] {
        if (!h) {
            return;
        }
        (void)hipStreamSynchronize(h->stream);
        for (void *p : h->pool) {
            (void)hipFree(p);
        }
        (void)hipFree(h->red);
        (void)hipFree(h->staging);
        for (auto &b : h->bufs) {
            (void)hipFree(b.first);
        }
        (void)hipFree(h->d);
        for (hipEvent_t e : h->slots) {
            if (e) {
                (void)hipEventDestroy(e);
            }
        }
        (void)hipEventDestroy(h->ev0);
        (void)hipEventDestroy(h->ev1);
        if (h->ev_order) {
            (void)hipEventDestroy(h->ev_order);
        }
        (void)hipStreamDestroy(h->stream);
        delete h;
    });
}

int pl_sv_num_qubits(pl_sv_t h, int *out) {
    return guarded([&] {
        if (!h || !out) throw InvalidArg("null argument");
        *out = h->n;
    });
}
int pl_sv_dtype(pl_sv_t h, int *out) {
    return guarded([&] {
        if (!h || !out) throw InvalidArg("null argument");
        *out = h->dtype;
    });
}

int pl_sv_upload(pl_sv_t h, const void *host, size_t nbytes) {
    return guarded([&] {
        if (!h || !host) throw InvalidArg("null argument");
        if (nbytes != h->state_bytes()) {
            throw InvalidArg("upload size does not match the state vector size");
        }
        discard_deferred(h);
        upload_host(h, h->d, host, nbytes);
        HIP_OK(hipStreamSynchronize(h->stream));
        g_h2d_bytes.fetch_add(nbytes, std::memory_order_relaxed);
    });
}
int pl_sv_download(pl_sv_t h, void *host, size_t nbytes) {
    return guarded([&] {
        if (!h || !host) throw InvalidArg("null argument");
        if (nbytes != h->state_bytes()) {
            throw InvalidArg("download size does not match the state vector size");
        }
        flush_deferred(h);
        download_host(h, host, h->d, nbytes);
        g_d2h_bytes.fetch_add(nbytes, std::memory_order_relaxed);
    });
}
int pl_transfer_counters(uint64_t *h2d_bytes, uint64_t *d2h_bytes) {
    return guarded([&] {
        if (h2d_bytes) *h2d_bytes = g_h2d_bytes.load(std::memory_order_relaxed);
        if (d2h_bytes) *d2h_bytes = g_d2h_bytes.load(std::memory_order_relaxed);
    });
}
int pl_sv_copy(pl_sv_t dst, pl_sv_t src) {
    return guarded([&] {
        if (!dst || !src) throw InvalidArg("null argument");
        if (dst->n != src->n || dst->dtype != src->dtype) {
            throw InvalidArg("state vectors differ in size or precision");
        }
        if (dst == src) {
            return;
        }
        flush_deferred(src);
        discard_deferred(dst);
        // dst's stream waits for src's queued work, copies, and src's stream
        // then waits for the copy, so later work queued on src (a gate) cannot
        // overwrite amplitudes the copy has not read yet
        HIP_OK(hipEventRecord(src->ev_order, src->stream));
        HIP_OK(hipStreamWaitEvent(dst->stream, src->ev_order, 0));
        HIP_OK(hipMemcpyAsync(dst->d, src->d, dst->state_bytes(), hipMemcpyDeviceToDevice, dst->stream));
        HIP_OK(hipEventRecord(dst->ev_order, dst->stream));
        HIP_OK(hipStreamWaitEvent(src->stream, dst->ev_order, 0));
    });
}
int pl_sv_set_basis_state(pl_sv_t h, uint64_t index) {
    return guarded([&] {
        if (!h) throw InvalidArg("null argument");
        if (index >= h->len) throw InvalidArg("basis index out of range");
        discard_deferred(h);
        const unsigned grid = grid_for(h->len);
        if (h->dtype == PL_C128) {
            hipLaunchKernelGGL((k_fill_basis<double>), dim3(grid), dim3(kBlock), 0, h->stream,
                               static_cast<double2 *>(h->d), h->len, index);
        } else {
            hipLaunchKernelGGL((k_fill_basis<float>), dim3(grid), dim3(kBlock), 0, h->stream,
                               static_cast<float2 *>(h->d), h->len, index);
        }
        HIP_OK(hipGetLastError());
    });
}
int pl_sv_reset(pl_sv_t h) { return pl_sv_set_basis_state(h, 0); }
int pl_sv_device_ptr(pl_sv_t h, void **out) {
    return guarded([&] {
        if (!h || !out) throw InvalidArg("null argument");
        flush_deferred(h); // the caller may read or write the memory itself
        *out = h->d;
    });
}
int pl_sv_synchronize(pl_sv_t h) {
    return guarded([&] {
        if (!h) throw InvalidArg("null argument");
        flush_deferred(h);
        HIP_OK(hipStreamSynchronize(h->stream));
    });
}

int pl_sv_set_deferred(pl_sv_t h, int on) {
    return guarded([&] {
        if (!h) throw InvalidArg("null argument");
        if (!on) {
            flush_deferred(h);
        }
        h->defer = on != 0;
    });
}
int pl_sv_deferred_count(pl_sv_t h, int *out) {
    return guarded([&] {
        if (!h || !out) throw InvalidArg("null argument");
        *out = static_cast<int>(h->q_ids.size());
    });
}

int pl_sv_apply_gate(pl_sv_t h, int id, const int *wires, int nw, int inverse,
                     const double *params, int np) {
    if (h && h->defer) {
        return guarded([&] { defer_gate(h, id, wires, nw, inverse != 0, params, np); });
    }
    return dispatch(h, [&](auto &E) { E.gate(E.psi(), id, wires, nw, inverse != 0, params, np); });
}
int pl_sv_apply_ops(pl_sv_t h, int nops, const int *ids, const int *nwires, const int *wires,
                    const uint8_t *inverse, const int *nparams, const double *params) {
    if (h && h->defer) {
        return guarded([&] {
            for (int i = 0, wo = 0, po = 0; i < nops; i++) {
                if (ids[i] != -1) {
                    defer_gate(h, ids[i], wires + wo, nwires[i], inverse[i] != 0, params + po, nparams[i]);
                }
                wo += nwires[i];
                po += nparams[i];
            }
        });
    }
    return dispatch(h, [&](auto &E) {
        for (int i = 0, wo = 0, po = 0; i < nops; i++) {
            if (ids[i] != -1) {
                E.gate(E.psi(), ids[i], wires + wo, nwires[i], inverse[i] != 0, params + po,
                       nparams[i]);
            }
            wo += nwires[i];
            po += nparams[i];
        }
    });
}

int pl_sv_apply_ops_fused(pl_sv_t h, int nops, const int *ids, const int *nwires, const int *wires,
                          const uint8_t *inverse, const int *nparams, const double *params,
                          int max_wires, int flags, int *nblocks) {
    return dispatch(h, [&](auto &E) {
        if (nops < 0 || (nops > 0 && (!ids || !nwires || !wires || !inverse || !nparams))) {
            throw InvalidArg("null tape argument");
        }
        const TapeView t{nops, ids, nwires, wires, inverse, nparams, params};
        const int nb = E.apply_fused(E.psi(), t, max_wires, flags, nullptr);
        if (nblocks) {
            *nblocks = nb;
        }
    });
}

int pl_sv_apply_ops_tiled(pl_sv_t h, int nops, const int *ids, const int *nwires, const int *wires,
                          const uint8_t *inverse, const int *nparams, const double *params,
                          int tile_bits, int low_bits, int flags, int *npasses) {
    return dispatch(h, [&](auto &E) {
        if (nops < 0 || (nops > 0 && (!ids || !nwires || !wires || !inverse || !nparams))) {
            throw InvalidArg("null tape argument");
        }
        if (flags & ~PL_FUSE_RENORM) {
            throw InvalidArg("unknown flag for the tiled executor");
        }
        const TapeView t{nops, ids, nwires, wires, inverse, nparams, params};
        const int np = E.apply_tiled(E.psi(), t, tile_bits, low_bits);
        if (flags & PL_FUSE_RENORM) {
            E.renormalize(E.psi());
        }
        if (npasses) {
            *npasses = np;
        }
    });
}

int pl_tile_plan(int num_qubits, int nops, const int *ids, const int *nwires, const int *wires,
                 const uint8_t *inverse, const int *nparams, const double *params, int tile_bits,
                 int low_bits, int capacity, int *npasses, int *pass_nops, int *pass_tiled,
                 int *pass_bits, int *order) {
    return guarded([&] {
        if (num_qubits < 1 || num_qubits > 62 || nops < 0 || !npasses) {
            throw InvalidArg("invalid argument");
        }
        if (nops > 0 && (!ids || !nwires || !wires || !inverse || !nparams)) {
            throw InvalidArg("null tape argument");
        }
        const TapeView t{nops, ids, nwires, wires, inverse, nparams, params};
        const auto passes = tile_schedule(num_qubits, t, tile_bits, low_bits);
        *npasses = static_cast<int>(passes.size());
        if (*npasses > capacity) {
            throw InvalidArg("capacity too small for " + std::to_string(*npasses) + " passes");
        }
        int at = 0;
        for (size_t p = 0; p < passes.size(); p++) {
            const auto &ps = passes[p];
            if (pass_nops) pass_nops[p] = static_cast<int>(ps.ops.size());
            if (pass_tiled) pass_tiled[p] = ps.tiled ? 1 : 0;
            for (int j = 0; pass_bits && j < 16; j++) {
                pass_bits[16 * p + j] = j < static_cast<int>(ps.hb.size()) ? ps.hb[j] : -1;
            }
            for (int i : ps.ops) {
                if (order) order[at] = i;
                at++;
            }
        }
    });
}

int pl_fuse_tape(int num_qubits, int nops, const int *ids, const int *nwires, const int *wires,
                 const uint8_t *inverse, const int *nparams, const double *params, int max_wires,
                 int pad_to, int capacity, int *nblocks, int *first_op, int *block_nops, int *order,
                 int *nbits, int *bits, double *mats) {
    return guarded([&] {
        if (num_qubits < 1 || num_qubits > 62 || nops < 0 || !nblocks) {
            throw InvalidArg("invalid argument");
        }
        const TapeView t{nops, ids, nwires, wires, inverse, nparams, params};
        const auto blocks = fuse_tape(num_qubits, t, max_wires, std::max(0, std::min(pad_to, 5)));
        *nblocks = static_cast<int>(blocks.size());
        if (*nblocks > capacity) {
            throw InvalidArg("capacity too small for " + std::to_string(*nblocks) + " blocks");
        }
        int at = 0;
        for (size_t b = 0; b < blocks.size(); b++) {
            const auto &B = blocks[b];
            if (first_op) first_op[b] = at;
            if (block_nops) block_nops[b] = static_cast<int>(B.ops.size());
            for (int i : B.ops) {
                if (order) order[at] = i;
                at++;
            }
            if (nbits) nbits[b] = B.fused ? static_cast<int>(B.bits.size()) : 0;
            for (size_t m = 0; bits && B.fused && m < B.bits.size(); m++) {
                bits[8 * b + m] = B.bits[m];
            }
            for (size_t j = 0; mats && B.fused && j < B.U.size(); j++) {
                mats[2048 * b + 2 * j] = B.U[j].real();
                mats[2048 * b + 2 * j + 1] = B.U[j].imag();
            }
        }
    });
}

int pl_sv_norm2(pl_sv_t h, double *out) {
    return dispatch(h, [&](auto &E) {
        if (!out) {
            throw InvalidArg("null output");
        }
        *out = E.dot(E.psi(), E.psi()).real();
    });
}

int pl_sv_distance(pl_sv_t a, pl_sv_t b, double *out) {
    return guarded([&] {
        if (!a || !b || !out) {
            throw InvalidArg("null argument");
        }
        if (a->n != b->n) {
            throw InvalidArg("states must have the same number of qubits");
        }
        HIP_OK(hipSetDevice(0));
        flush_deferred(a);
        flush_deferred(b);
        HIP_OK(hipStreamSynchronize(b->stream)); // b's pending work, then run on a's stream
        const unsigned grid = std::min<unsigned>(grid_for(a->len, 1, 4), kReduceBlocks);
        auto *part = reinterpret_cast<double2 *>(reduce_buf(a, (kReduceBlocks + 1) * sizeof(double2)));
        auto launch = [&](auto *pa, auto *pb) {
            using A = std::remove_pointer_t<decltype(pa)>;
            using B = std::remove_pointer_t<decltype(pb)>;
            hipLaunchKernelGGL((k_diff_norm<typename A::value_type, typename B::value_type>), dim3(grid),
                               dim3(kBlock), 0, a->stream, reinterpret_cast<const typename A::type *>(a->d),
                               reinterpret_cast<const typename B::type *>(b->d), a->len, part);
        };
        struct D { using value_type = double; using type = v2<double>; };
        struct F { using value_type = float; using type = v2<float>; };
        if (a->dtype == PL_C128 && b->dtype == PL_C128) launch((D *)nullptr, (D *)nullptr);
        else if (a->dtype == PL_C128) launch((D *)nullptr, (F *)nullptr);
        else if (b->dtype == PL_C128) launch((F *)nullptr, (D *)nullptr);
        else launch((F *)nullptr, (F *)nullptr);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, a->stream, part, (int)grid,
                           part + kReduceBlocks);
        HIP_OK(hipGetLastError());
        double2 r;
        HIP_OK(hipMemcpyAsync(&r, part + kReduceBlocks, sizeof(r), hipMemcpyDeviceToHost, a->stream));
        HIP_OK(hipStreamSynchronize(a->stream));
        *out = std::sqrt(r.x);
    });
}

int pl_sv_apply_matrix(pl_sv_t h, const double *m, const int *wires, int nw, int inverse) {
    return dispatch(h, [&](auto &E) { E.matrix(E.psi(), m, wires, nw, inverse != 0); });
}

int pl_sv_apply_generator(pl_sv_t h, int id, const int *wires, int nw, int adj, double *scale) {
    (void)adj; // the generators ignore `adj` (GateImplementationsLM.hpp:1355-1640)
    return dispatch(h, [&](auto &E) {
        const double s = E.generator(E.psi(), id, wires, nw);
        if (scale) {
            *scale = s;
        }
    });
}

int pl_sv_apply_observable(pl_sv_t h, const int *prog, const double *coeffs, const double *mats) {
    return dispatch(h, [&](auto &E) {
        if (!prog) throw InvalidArg("null observable program");
        ObsReader rd{prog + 1, coeffs, mats};
        E.apply_obs(E.psi(), rd.one());
    });
}

int pl_expval(pl_sv_t h, const int *prog, const double *coeffs, const double *mats, double *out) {
    return dispatch(h, [&](auto &E) {
        ObsReader rd{prog + 1, coeffs, mats};
        *out = E.expval(rd.one(), false);
    });
}
// ---- allocation-free expectation value --------------------------------------
// Device scratch of pl_expval_noalloc: the term tables and partial sums of
// one call.  Each GPU the library runs on holds its own copy until the
// process ends; every call on that GPU reuses it, so calls that overlap on
// one GPU race.  Nothing in it needs clearing: each call writes every
// partial it reads.
constexpr size_t kNoAllocScratchBytes = size_t{2} << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is not a device address).
unsigned char *noalloc_scratch() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        throw DeviceError("device index out of range for the allocation-free scratch");
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    return g_scratch_ptr[dev];
}
} // namespace

int pl_expval_noalloc(pl_sv_t h, const int *prog, const double *coeffs, const double *mats, double *out) {
    if (h && !h->q_ids.empty()) { // applying queued gates may allocate
        return guarded([] {
            throw InvalidArg("gates are queued on this state (pl_sv_set_deferred): call pl_sv_synchronize "
                             "before the allocation-free expectation value");
        });
    }
    return dispatch(h, [&](auto &E) {
        if (!prog || !out) {
            throw InvalidArg("null argument");
        }
        ObsReader rd{prog + 1, coeffs, mats};
        const ObsNode o = rd.one();
        std::vector<PauliWord> ws;
        std::vector<double> cs;
        if (!E.as_pauli_sum(o, ws, cs)) {
            throw InvalidArg("the allocation-free expectation value takes Pauli words and Hamiltonians of "
                             "Pauli words only (other observables need a state-sized scratch)");
        }
        *out = E.pauli_expval(E.psi(), ws, cs, noalloc_scratch(), kNoAllocScratchBytes);
    });
}

int pl_var(pl_sv_t h, const int *prog, const double *coeffs, const double *mats, double *out) {
    return dispatch(h, [&](auto &E) {
        ObsReader rd{prog + 1, coeffs, mats};
        *out = E.expval(rd.one(), true);
    });
}

int pl_expval_named(pl_sv_t h, int id, const int *wires, int nw, double *out) {
    return dispatch(h, [&](auto &E) {
        ObsNode o;
        o.kind = 0;
        o.op = id;
        gate_name(id);
        o.wires.assign(wires, wires + nw);
        *out = E.expval(o, false);
    });
}
int pl_var_named(pl_sv_t h, int id, const int *wires, int nw, double *out) {
    return dispatch(h, [&](auto &E) {
        ObsNode o;
        o.kind = 0;
        o.op = id;
        gate_name(id);
        o.wires.assign(wires, wires + nw);
        *out = E.expval(o, true);
    });
}
int pl_expval_matrix(pl_sv_t h, const double *m, const int *wires, int nw, double *out) {
    return dispatch(h, [&](auto &E) {
        if (nw <= 0 || nw > kMaxPos) {
            throw InvalidArg("matrix observable needs 1..12 wires");
        }
        ObsNode o;
        o.kind = 1;
        o.wires.assign(wires, wires + nw);
        const size_t dim = size_t{1} << (2 * nw);
        for (size_t i = 0; i < dim; i++) {
            o.mat.emplace_back(m[2 * i], m[2 * i + 1]);
        }
        *out = E.expval(o, false);
    });
}

int pl_probs(pl_sv_t h, const int *wires, int nw, double *out) {
    return dispatch(h, [&](auto &E) {
        if (!out) throw InvalidArg("null output");
        E.probs(wires, nw, out);
    });
}

int pl_sample(pl_sv_t h, const double *uniforms, int64_t m, int64_t *idx) {
    return dispatch(h, [&](auto &E) {
        if (m < 0) throw InvalidArg("negative sample count");
        if (m == 0) return;
        if (!uniforms || !idx) throw InvalidArg("null argument");
        E.sample(uniforms, 0, m, idx, nullptr);
    });
}

int pl_generate_samples(pl_sv_t h, uint64_t seed, int64_t m, uint64_t *bits) {
    return dispatch(h, [&](auto &E) {
        if (m < 0) throw InvalidArg("negative sample count");
        if (m == 0) return;
        if (!bits) throw InvalidArg("null argument");
        E.sample(nullptr, seed, m, nullptr, bits);
    });
}

int pl_adjoint_jacobian(pl_sv_t h, int nops, const int *ids, const int *nwires, const int *wires,
                        const uint8_t *inverse, const int *nparams, const double *params,
                        const int *prog, const double *coeffs, const double *mats, int ntp,
                        const int *tp, double *jac) {
    return dispatch(h, [&](auto &E) {
        adjoint_sweep(E, nops, ids, nwires, wires, inverse, nparams, params, prog, coeffs, mats,
                      ntp, tp, jac);
    });
}

int pl_vjp(pl_sv_t h, int nops, const int *ids, const int *nwires, const int *wires,
           const uint8_t *inverse, const int *nparams, const double *params, const double *dy,
           int ntp, const int *tp, double *vjp) {
    return dispatch(h, [&](auto &E) {
        if (!dy || (ntp > 0 && (!tp || !vjp)) || nops < 0) {
            throw InvalidArg("null argument");
        }
        vjp_sweep(E, nops, ids, nwires, wires, inverse, nparams, params, dy, ntp, tp, vjp);
    });
}

// Measurements::expval / var of a CSR matrix (MeasurementsLQubit.hpp:183-214,
// 437-471; SparseLinAlg.hpp apply_Sparse_Matrix): both from one fused pass.
static void sparse_moments(pl_sv_t h, const int64_t *row_map, int64_t row_map_size,
                           const int64_t *entries, const double *values, int64_t nnz, double *mean,
                           double *mean_square) {
    if (!row_map || (nnz > 0 && (!entries || !values)) || nnz < 0) {
        throw InvalidArg("null argument");
    }
    if (static_cast<uint64_t>(row_map_size) != h->len + 1) {
        throw InvalidArg("Statevector and Hamiltonian have incompatible sizes.");
    }
    if (row_map[0] != 0 || row_map[row_map_size - 1] != nnz) {
        throw InvalidArg("row_map must start at 0 and end at the number of non-zeros");
    }
    // the order of row_map and the range of every column index are checked
    // by k_sparse_moments as it reads them (one pass over the arrays, on the
    // device); the result is discarded and the call fails if either is wrong
    dispatch_or_throw(h, [&](auto &E) {
        using T = typename std::remove_reference_t<decltype(E)>::Scalar;
        // complex128 values are uploaded as given; complex64 ones are rounded first
        std::vector<dcx<T>> vals;
        const void *vsrc = values;
        if constexpr (!std::is_same_v<T, double>) {
            vals.resize(std::max<int64_t>(nnz, 1));
            for (int64_t j = 0; j < nnz; j++) {
                vals[j] = {static_cast<T>(values[2 * j]), static_cast<T>(values[2 * j + 1])};
            }
            vsrc = vals.data();
        }
        unsigned char *dev = nullptr;
        const size_t rb = row_map_size * 8, eb = std::max<int64_t>(nnz, 1) * 8,
                     vb = std::max<int64_t>(nnz, 1) * sizeof(dcx<T>);
        HIP_OK(hipMalloc(&dev, rb + eb + vb));
        struct Free {
            unsigned char *p;
            ~Free() { (void)hipFree(p); }
        } fr{dev};
        upload_host(h, dev, row_map, rb);
        if (nnz > 0) {
            upload_host(h, dev + rb, entries, nnz * 8);
            upload_host(h, dev + rb + eb, vsrc, nnz * sizeof(dcx<T>));
        }
        const unsigned grid = std::min<unsigned>(grid_for(E.len(), 1, 4), kReduceBlocks);
        // partials, then the result, then the check flags (read back together)
        auto *part = reinterpret_cast<double2 *>(reduce_buf(h, (kReduceBlocks + 2) * sizeof(double2)));
        auto *bad = reinterpret_cast<unsigned *>(part + kReduceBlocks + 1);
        HIP_OK(hipMemsetAsync(bad, 0, sizeof(unsigned), h->stream));
        hipLaunchKernelGGL((k_sparse_moments<T>), dim3(grid), dim3(kBlock), 0, h->stream, E.psi(), E.len(),
                           reinterpret_cast<const int64_t *>(dev), reinterpret_cast<const int64_t *>(dev + rb),
                           reinterpret_cast<const dcx<T> *>(dev + rb + eb), nnz, bad, part);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, h->stream, part, (int)grid,
                           part + kReduceBlocks);
        HIP_OK(hipGetLastError());
        double2 r[2];
        HIP_OK(hipMemcpyAsync(r, part + kReduceBlocks, sizeof(r), hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        unsigned flags = 0;
        std::memcpy(&flags, &r[1], sizeof(flags));
        if (flags & 1U) {
            throw InvalidArg("row_map must be non-decreasing");
        }
        if (flags & 2U) {
            throw InvalidArg("column index out of range");
        }
        *mean = r[0].x;
        *mean_square = r[0].y;
    });
}

int pl_expval_sparse(pl_sv_t h, const int64_t *row_map, int64_t row_map_size, const int64_t *entries,
                     const double *values, int64_t nnz, double *out) {
    return guarded([&] {
        if (!h || !out) throw InvalidArg("null argument");
        double m = 0, ms = 0;
        sparse_moments(h, row_map, row_map_size, entries, values, nnz, &m, &ms);
        *out = m;
    });
}

int pl_var_sparse(pl_sv_t h, const int64_t *row_map, int64_t row_map_size, const int64_t *entries,
                  const double *values, int64_t nnz, double *out) {
    return guarded([&] {
        if (!h || !out) throw InvalidArg("null argument");
        double m = 0, ms = 0;
        sparse_moments(h, row_map, row_map_size, entries, values, nnz, &m, &ms);
        *out = ms - m * m;
    });
}

int pl_generate_mcmc_samples(pl_sv_t h, int kernel, int64_t num_burnin, int64_t nshots, uint64_t seed,
                             int nchains, uint64_t *bits) {
    return dispatch(h, [&](auto &E) {
        if (!bits) {
            throw InvalidArg("null output");
        }
        E.mcmc(kernel, num_burnin, nshots, seed, nchains, bits);
    });
}

int pl_timer_start(pl_sv_t h) {
    return guarded([&] {
        if (!h) throw InvalidArg("null argument");
        flush_deferred(h);
        HIP_OK(hipEventRecord(h->ev0, h->stream));
    });
}
int pl_timer_stop(pl_sv_t h, double *ms) {
    return guarded([&] {
        if (!h || !ms) throw InvalidArg("null argument");
        flush_deferred(h);
        HIP_OK(hipEventRecord(h->ev1, h->stream));
        HIP_OK(hipEventSynchronize(h->ev1));
        float f = 0;
        HIP_OK(hipEventElapsedTime(&f, h->ev0, h->ev1));
        *ms = f;
    });
}

int pl_event_record(pl_sv_t h, int slot) {
    return guarded([&] {
        if (!h) throw InvalidArg("null argument");
        if (slot < 0 || slot >= 65536) throw InvalidArg("event slot out of range");
        if (h->slots.size() <= static_cast<size_t>(slot)) {
            h->slots.resize(slot + 1, nullptr);
        }
        if (!h->slots[slot]) {
            HIP_OK(hipEventCreate(&h->slots[slot]));
        }
        flush_deferred(h);
        HIP_OK(hipEventRecord(h->slots[slot], h->stream));
    });
}
int pl_event_elapsed(pl_sv_t h, int a, int b, double *ms) {
    return guarded([&] {
        if (!h || !ms) throw InvalidArg("null argument");
        if (a < 0 || b < 0 || static_cast<size_t>(std::max(a, b)) >= h->slots.size() ||
            !h->slots[a] || !h->slots[b]) {
            throw InvalidArg("event slot was never recorded");
        }
        HIP_OK(hipEventSynchronize(h->slots[b]));
        float f = 0;
        HIP_OK(hipEventElapsedTime(&f, h->slots[a], h->slots[b]));
        *ms = f;
    });
}

} // extern "C"
