// Host side of the batched layered-ansatz engine (kernels_batch.hpp):
// circuit tables, entangler maps over GF(2), per-launch block tables, the
// tile-width autotuner and the hipGraph that replays one whole batch
// (forward, observable, adjoint) as a single launch.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels_batch.hpp"

namespace plamd {

struct BatchPlanBase {
    virtual ~BatchPlanBase() = default;
    virtual void run(const double *params, double *energies, double *grads) = 0;
    virtual int tile_width() const = 0;
    virtual int launches() const = 0;
    virtual uint64_t total_params() const = 0;
};

// y = M x over GF(2), x and y bit vectors in bit-position order (bit b of the
// basis index = wire n-1-b).  CNOT(c, t) does x_t ^= x_c.
struct Gf2Map {
    int n;
    uint32_t row[kBatchMaxQubits]; // y_b = parity(x & row[b])
    explicit Gf2Map(int nn) : n(nn), row{} {
        for (int b = 0; b < n; b++) {
            row[b] = 1U << b;
        }
    }
    void cnot(int cw, int tw) { // wires
        const int cb = n - 1 - cw, tb = n - 1 - tw;
        row[tb] ^= row[cb];
    }
    Gf2Map inverse() const {
        // Gauss-Jordan on [M | I]
        uint32_t a[kBatchMaxQubits], inv[kBatchMaxQubits];
        for (int b = 0; b < n; b++) {
            a[b] = row[b];
            inv[b] = 1U << b;
        }
        for (int col = 0; col < n; col++) {
            int piv = col;
            while (piv < n && !((a[piv] >> col) & 1U)) {
                piv++;
            }
            if (piv == n) {
                throw InvalidArg("entangler map is singular");
            }
            std::swap(a[piv], a[col]);
            std::swap(inv[piv], inv[col]);
            for (int r = 0; r < n; r++) {
                if (r != col && ((a[r] >> col) & 1U)) {
                    a[r] ^= a[col];
                    inv[r] ^= inv[col];
                }
            }
        }
        Gf2Map m(n);
        for (int b = 0; b < n; b++) {
            m.row[b] = inv[b];
        }
        return m;
    }
};

template <class T> struct BatchPlan final : BatchPlanBase {
    using V = v2<T>;
    struct LaunchSpec {
        enum Kind { Init, Rot, Ring, Obs, RotAdj } kind;
        int layer = 0, stage = 0;
        BLaunch L{};
        uint64_t blocks = 0;
    };

    int nc = 0, npat = 0, entangler = 0;
    int4 pat{0, 0, 0, 0};
    std::vector<int> nq, depth;
    std::vector<PauliTerm> terms;
    std::vector<uint64_t> xmasks;
    std::vector<uint32_t> term_off, term_cnt;
    std::vector<uint64_t> par_off;
    uint64_t npar = 0;
    std::vector<double> norms; // <psi|psi> of each circuit's final state, last replay

    int W = 0;
    hipStream_t stream = nullptr;
    V *arena = nullptr;
    BCirc *dcircs = nullptr;
    PauliTerm *dterms = nullptr;
    uint64_t *dxm = nullptr;
    double *dparams = nullptr, *dgrads = nullptr, *denergy = nullptr;
    double2 *dcs = nullptr; // (cos, sin) of every half angle, refreshed each replay
    int *dcirc_lists = nullptr;
    uint64_t *dprefix = nullptr;
    std::vector<LaunchSpec> seq;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    BatchPlan(int ncirc, const int *n, const int *d, const char *pattern, int ent,
              const std::vector<std::vector<PauliTerm>> &t, const std::vector<std::vector<uint64_t>> &x,
              int tile_width) {
        nc = ncirc;
        entangler = ent;
        npat = static_cast<int>(std::strlen(pattern));
        if (npat < 1 || npat > kBatchMaxPattern) {
            throw InvalidArg("rotation pattern must have 1..4 letters");
        }
        int ax[4] = {0, 0, 0, 0};
        for (int a = 0; a < npat; a++) {
            const char ch = pattern[a];
            if (ch != 'X' && ch != 'Y' && ch != 'Z') {
                throw InvalidArg("rotation pattern letters must be X, Y or Z");
            }
            ax[a] = ch - 'X';
        }
        pat = int4{ax[0], ax[1], ax[2], ax[3]};
        if (ent < 0 || ent > 2) {
            throw InvalidArg("entangler must be 0 (none), 1 (CNOT ladder) or 2 (CNOT ring)");
        }
        for (int j = 0; j < nc; j++) {
            if (n[j] < 1 || n[j] > kBatchMaxQubits || d[j] < 0) {
                throw InvalidArg("circuit qubit counts must be in [1, 30] and depths >= 0");
            }
            nq.push_back(n[j]);
            depth.push_back(d[j]);
            par_off.push_back(npar);
            npar += static_cast<uint64_t>(d[j]) * n[j] * npat;
            term_off.push_back(static_cast<uint32_t>(terms.size()));
            term_cnt.push_back(static_cast<uint32_t>(t[j].size()));
            // group the circuit's terms by x-mask: one amplitude load per mask
            std::vector<size_t> order(t[j].size());
            std::iota(order.begin(), order.end(), size_t{0});
            std::stable_sort(order.begin(), order.end(),
                             [&](size_t a, size_t b) { return x[j][a] < x[j][b]; });
            for (size_t k : order) {
                terms.push_back(t[j][k]);
                xmasks.push_back(x[j][k]);
            }
        }
        int dev = 0;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        if (tile_width != 0 && (tile_width < 4 || tile_width > max_tile())) {
            throw InvalidArg("tile width must be 0 (autotune) or in [4, " + std::to_string(max_tile()) + "]");
        }
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        uint64_t amps = 0;
        for (int j = 0; j < nc; j++) {
            amps += uint64_t{4} << nq[j];
        }
        HIP_OK(hipMalloc(&arena, std::max<uint64_t>(amps, 1) * sizeof(V)));
        HIP_OK(hipMalloc(&dcircs, std::max(nc, 1) * sizeof(BCirc)));
        HIP_OK(hipMalloc(&dterms, std::max<size_t>(terms.size(), 1) * sizeof(PauliTerm)));
        HIP_OK(hipMalloc(&dxm, std::max<size_t>(xmasks.size(), 1) * sizeof(uint64_t)));
        HIP_OK(hipMalloc(&dparams, std::max<uint64_t>(npar, 1) * sizeof(double)));
        HIP_OK(hipMalloc(&dgrads, std::max<uint64_t>(npar, 1) * sizeof(double)));
        HIP_OK(hipMalloc(&dcs, std::max<uint64_t>(npar, 1) * sizeof(double2)));
        HIP_OK(hipMalloc(&denergy, 2 * std::max(nc, 1) * sizeof(double))); // energies, then <psi|psi>
        if (!terms.empty()) {
            HIP_OK(hipMemcpy(dterms, terms.data(), terms.size() * sizeof(PauliTerm), hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(dxm, xmasks.data(), xmasks.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        }
        if (tile_width > 0) {
            configure(tile_width);
        } else {
            autotune();
        }
    }

    ~BatchPlan() override {
        if (exec) {
            (void)hipGraphExecDestroy(exec);
        }
        if (graph) {
            (void)hipGraphDestroy(graph);
        }
        for (void *p : {static_cast<void *>(arena), static_cast<void *>(dcircs),
                        static_cast<void *>(dterms), static_cast<void *>(dxm),
                        static_cast<void *>(dparams), static_cast<void *>(dgrads), static_cast<void *>(dcs),
                        static_cast<void *>(denergy), static_cast<void *>(dcirc_lists),
                        static_cast<void *>(dprefix)}) {
            if (p) {
                (void)hipFree(p);
            }
        }
        if (stream) {
            (void)hipStreamDestroy(stream);
        }
    }

    int lds_limit = 65536; // bytes of LDS a workgroup may use (device attribute)

    // widest tile whose adjoint stage (two tiles + its 2 KiB of reduction slots) fits
    int max_tile() const {
        int w = 4;
        while (w < kBatchMaxTile && (size_t{2} << (w + 1)) * sizeof(V) + 2048 <= static_cast<size_t>(lds_limit)) {
            w++;
        }
        return w;
    }

    static int stages(int n, int w) { return n <= w ? 1 : 1 + (n - w + (w - 4)) / (w - 3); }
    // tile width of stage s (stage_bits in kernels_batch.hpp): later stages
    // always use full W-bit tiles
    static int stage_T(int n, int w, int s) { return s == 0 ? std::min(n, w) : w; }

    // Build circuit tables and the launch sequence for tile width w.
    void configure(int w) {
        if (w < 4 || w > max_tile()) {
            throw InvalidArg("tile width out of range");
        }
        W = w;
        std::vector<BCirc> hc(nc);
        uint64_t off = 0;
        for (int j = 0; j < nc; j++) {
            BCirc &c = hc[j];
            std::memset(&c, 0, sizeof(c));
            c.n = nq[j];
            c.depth = depth[j];
            c.nstage = stages(nq[j], W);
            c.nterms = static_cast<int>(term_cnt[j]);
            c.off = off;
            off += uint64_t{4} << nq[j];
            c.par = static_cast<uint32_t>(par_off[j]);
            c.term = term_off[j];
            Gf2Map m(nq[j]);
            if (entangler && nq[j] > 1) {
                const int last = entangler == 2 ? nq[j] : nq[j] - 1;
                for (int i = 0; i < last; i++) {
                    m.cnot(i, (i + 1) % nq[j]);
                }
            }
            const Gf2Map mi = m.inverse();
            for (int b = 0; b < nq[j]; b++) {
                c.fwd_rows[b] = mi.row[b]; // out[y] = in[M^-1 y]
                c.adj_rows[b] = m.row[b];  // out[x] = in[M x]
            }
        }
        HIP_OK(hipMemcpy(dcircs, hc.data(), nc * sizeof(BCirc), hipMemcpyHostToDevice));

        // launch tables
        seq.clear();
        std::vector<int> lists;
        std::vector<uint64_t> prefixes;
        std::vector<std::pair<size_t, size_t>> where; // (list offset, prefix offset)
        auto add = [&](typename LaunchSpec::Kind kind, int layer, int stage, auto active, auto blocks) {
            LaunchSpec ls;
            ls.kind = kind;
            ls.layer = layer;
            ls.stage = stage;
            const size_t lo = lists.size(), po = prefixes.size();
            uint64_t tot = 0;
            int cnt = 0;
            for (int j = 0; j < nc; j++) {
                if (active(j)) {
                    lists.push_back(j);
                    prefixes.push_back(tot);
                    tot += blocks(j);
                    cnt++;
                }
            }
            prefixes.push_back(tot);
            if (tot * kBlock > 0xFFFFFFFFull) { // work-items per launch dimension
                throw InvalidArg("batch too large for tile width " + std::to_string(w) +
                                 ": one launch would exceed 2^32 work-items");
            }
            if (cnt == 0) {
                lists.resize(lo);
                prefixes.resize(po);
                return;
            }
            ls.L.count = cnt;
            ls.blocks = tot;
            where.emplace_back(lo, po);
            seq.push_back(ls);
        };
        const int dmax = nc ? *std::max_element(depth.begin(), depth.end()) : 0;
        int smax = 1;
        for (int j = 0; j < nc; j++) {
            smax = std::max(smax, stages(nq[j], W));
        }
        auto elem_blocks = [&](int j) { return ((uint64_t{1} << nq[j]) + kBlock - 1) / kBlock; };
        add(LaunchSpec::Init, 0, 0, [&](int) { return true; }, elem_blocks);
        for (int l = 0; l < dmax; l++) {
            for (int s = 0; s < smax; s++) {
                add(LaunchSpec::Rot, l, s,
                    [&](int j) { return depth[j] > l && stages(nq[j], W) > s; },
                    [&](int j) { return uint64_t{1} << (nq[j] - stage_T(nq[j], W, s)); });
            }
            // a layer's entangler is fused into the next layer's first
            // rotation stage; only a circuit's last one runs on its own
            if (entangler) {
                add(LaunchSpec::Ring, l, 0, [&](int j) { return depth[j] == l + 1; }, elem_blocks);
            }
        }
        add(LaunchSpec::Obs, 0, 0, [&](int) { return true; }, elem_blocks);
        // the entangler of each layer is undone inside its first adjoint
        // rotation stage (k_batch_rot_adj gathers lambda and H)
        for (int l = dmax - 1; l >= 0; l--) {
            for (int s = smax - 1; s >= 0; s--) {
                add(LaunchSpec::RotAdj, l, s,
                    [&](int j) { return depth[j] > l && stages(nq[j], W) > s; },
                    [&](int j) { return uint64_t{1} << (nq[j] - stage_T(nq[j], W, s)); });
            }
        }
        if (dcirc_lists) {
            HIP_OK(hipFree(dcirc_lists));
            dcirc_lists = nullptr;
        }
        if (dprefix) {
            HIP_OK(hipFree(dprefix));
            dprefix = nullptr;
        }
        HIP_OK(hipMalloc(&dcirc_lists, std::max<size_t>(lists.size(), 1) * sizeof(int)));
        HIP_OK(hipMalloc(&dprefix, std::max<size_t>(prefixes.size(), 1) * sizeof(uint64_t)));
        HIP_OK(hipMemcpy(dcirc_lists, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dprefix, prefixes.data(), prefixes.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        for (size_t i = 0; i < seq.size(); i++) {
            seq[i].L.circ = dcirc_lists + where[i].first;
            seq[i].L.prefix = dprefix + where[i].second;
        }
        capture();
    }

    void enqueue() {
        const size_t tile_bytes = (size_t{1} << W) * sizeof(V);
        hipLaunchKernelGGL(k_zero, dim3((npar + kBlock - 1) / kBlock + 1), dim3(kBlock), 0, stream,
                           dgrads, npar);
        hipLaunchKernelGGL(k_zero, dim3((2 * nc + kBlock - 1) / kBlock + 1), dim3(kBlock), 0, stream,
                           denergy, static_cast<uint64_t>(2 * nc));
        hipLaunchKernelGGL(k_batch_trig, dim3((npar + kBlock - 1) / kBlock + 1), dim3(kBlock), 0, stream,
                           dparams, npar, dcs);
        for (const auto &ls : seq) {
            const dim3 grid(static_cast<unsigned>(ls.blocks));
            const int cur = entangler ? (ls.layer & 1) : 0;
            const int nxt = entangler ? 1 - cur : 0;
            switch (ls.kind) {
            case LaunchSpec::Init:
                hipLaunchKernelGGL((k_batch_init<T>), grid, dim3(kBlock), 0, stream, arena, dcircs, ls.L);
                break;
            case LaunchSpec::Rot: {
                // stage 0 of layer l >= 1 gathers through layer l-1's entangler
                const int gsrc = (entangler && ls.layer > 0 && ls.stage == 0) ? nxt : -1;
                hipLaunchKernelGGL((k_batch_rot<T>), grid, dim3(kBlock), tile_bytes, stream, arena,
                                   dcircs, ls.L, W, ls.stage, ls.layer, cur, gsrc, npat, pat, dcs);
                break;
            }
            case LaunchSpec::Ring:
                hipLaunchKernelGGL((k_batch_ring<T>), grid, dim3(kBlock), 0, stream, arena, dcircs,
                                   ls.L, 0, cur, nxt, -1, -1);
                break;
            case LaunchSpec::Obs:
                hipLaunchKernelGGL((k_batch_obs<T>), grid, dim3(kBlock), 0, stream, arena, dcircs,
                                   ls.L, entangler ? 1 : 0, dterms, dxm, denergy,
                                   denergy + nc);
                break;
            case LaunchSpec::RotAdj:
                // lambda and H of layer l+1's side (buffers nxt) are gathered
                // through the inverse entangler at each circuit's first stage
                hipLaunchKernelGGL((k_batch_rot_adj<T>), grid, dim3(kBlock), 2 * tile_bytes, stream,
                                   arena, dcircs, ls.L, W, ls.stage, ls.layer, cur, 2 + cur,
                                   entangler ? nxt : -1, entangler ? 2 + nxt : -1, npat, pat, dcs, dgrads);
                break;
            }
            const hipError_t err = hipGetLastError();
            if (err != hipSuccess) {
                throw DeviceError("batch launch (kind " + std::to_string(static_cast<int>(ls.kind)) +
                                  ", layer " + std::to_string(ls.layer) + ", stage " +
                                  std::to_string(ls.stage) + ", " + std::to_string(ls.blocks) +
                                  " blocks, W " + std::to_string(W) + "): " + hipGetErrorString(err));
            }
        }
    }

    void capture() {
        if (exec) {
            HIP_OK(hipGraphExecDestroy(exec));
            exec = nullptr;
        }
        if (graph) {
            HIP_OK(hipGraphDestroy(graph));
            graph = nullptr;
        }
        HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        enqueue();
        HIP_OK(hipStreamEndCapture(stream, &graph));
        HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }

    // Project autotuner: time one graph replay per candidate tile width on
    // the plan's own circuits (random angles) and keep the fastest.
    void autotune() {
        std::vector<double> p(npar);
        for (uint64_t i = 0; i < npar; i++) {
            p[i] = 0.1 + 0.37 * static_cast<double>(i % 17);
        }
        std::vector<double> e(nc), g(npar);
        int best = max_tile();
        double best_t = 1e300;
        int nmax = 1;
        for (int j = 0; j < nc; j++) {
            nmax = std::max(nmax, nq[j]);
        }
        for (int w = 8; w <= max_tile(); w++) {
            if (w > nmax && w != 8) {
                break;
            }
            try {
                configure(w);
            } catch (const InvalidArg &) { // this width needs too many work-items
                continue;
            }
            run(p.data(), e.data(), g.data()); // warm-up
            const auto t0 = std::chrono::steady_clock::now();
            run(p.data(), e.data(), g.data());
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (dt < best_t) {
                best_t = dt;
                best = w;
            }
        }
        configure(best);
    }

    void run(const double *params, double *energies, double *grads) override {
        HIP_OK(hipMemcpyAsync(dparams, params, npar * sizeof(double), hipMemcpyHostToDevice, stream));
        HIP_OK(hipGraphLaunch(exec, stream));
        norms.resize(nc);
        HIP_OK(hipMemcpyAsync(energies, denergy, nc * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(norms.data(), denergy + nc, nc * sizeof(double), hipMemcpyDeviceToHost, stream));
        if (grads) {
            HIP_OK(hipMemcpyAsync(grads, dgrads, npar * sizeof(double), hipMemcpyDeviceToHost, stream));
        }
        HIP_OK(hipStreamSynchronize(stream));
        if constexpr (!std::is_same_v<T, double>) {
            // complex64: the FP64 norm of each final state removes the norm
            // drift of FP32 rounding from E = <psi|H|psi> / <psi|psi> and from
            // its gradient (d<psi|psi>/dtheta = 0 for a unitary circuit)
            uint64_t k = 0;
            for (int j = 0; j < nc; j++) {
                const double inv = norms[j] > 0 ? 1.0 / norms[j] : 1.0;
                energies[j] *= inv;
                const uint64_t m = static_cast<uint64_t>(depth[j]) * nq[j] * npat;
                if (grads) {
                    for (uint64_t a = 0; a < m; a++) {
                        grads[k + a] *= inv;
                    }
                }
                k += m;
            }
        }
    }
    int tile_width() const override { return W; }
    int launches() const override { return static_cast<int>(seq.size()) + 3; }
    uint64_t total_params() const override { return npar; }
};

} // namespace plamd
