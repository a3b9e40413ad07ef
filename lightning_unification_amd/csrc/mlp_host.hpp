// Host side of the learned surrogate: a d0 -> d1 -> d2 -> d3 MLP (GELU on the
// two hidden layers) whose weights live in HBM as FP32 masters plus BF16
// copies for the matrix cores.  Forward, backward and Adam run as the kernels
// of kernels_mlp.hpp on one stream; training data stays resident on the
// device (features in BF16, targets in FP32) and minibatches are contiguous
// slices of it.
//
// Parameter layout (host and device, FP32): W1 [d0 x d1] row-major, b1 [d1],
// W2 [d1 x d2], b2 [d2], W3 [d2 x d3], b3 [d3].  y = act2 W3 + b3 with
// act_k = GELU(act_{k-1} W_k + b_k), act_0 = x.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "kernels_mlp.hpp"

namespace plamd {

struct Mlp {
    int d[4];
    int maxb;
    size_t woff[3], boff[3], nparam = 0;
    hipStream_t stream = nullptr;
    float *p32 = nullptr, *grad = nullptr, *mom = nullptr, *vel = nullptr; // nparam each
    uint16_t *pbf = nullptr;                                              // BF16 copy
    uint16_t *xb = nullptr, *a1 = nullptr, *a2 = nullptr;                 // maxb x d0 / d1 / d2
    float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *dy = nullptr, *dz1 = nullptr, *dz2 = nullptr;
    float *tb = nullptr, *colw = nullptr, *xf = nullptr;
    double *lpart = nullptr;
    uint16_t *data_x = nullptr;
    float *data_t = nullptr;
    int64_t ndata = 0;
    int64_t step = 0, cursor = 0;
    static constexpr int kLossBlocks = 256;

    Mlp(int d0, int d1, int d2, int d3, int max_batch, uint64_t seed) : d{d0, d1, d2, d3}, maxb(max_batch) {
        for (int l = 0; l < 3; l++) {
            woff[l] = nparam;
            nparam += static_cast<size_t>(d[l]) * d[l + 1];
            boff[l] = nparam;
            nparam += d[l + 1];
        }
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        auto alloc = [](auto **p, size_t count) {
            HIP_OK(hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(**p)));
        };
        alloc(&p32, nparam);
        alloc(&grad, nparam);
        alloc(&mom, nparam);
        alloc(&vel, nparam);
        alloc(&pbf, nparam);
        const size_t B = maxb;
        alloc(&xf, B * d0);
        alloc(&xb, B * d0);
        alloc(&a1, B * d1);
        alloc(&a2, B * d2);
        alloc(&z1, B * d1);
        alloc(&z2, B * d2);
        alloc(&dz1, B * d1);
        alloc(&dz2, B * d2);
        alloc(&y, B * d3);
        alloc(&dy, B * d3);
        alloc(&tb, B * d3);
        alloc(&colw, d3);
        alloc(&lpart, kLossBlocks);
        // He-uniform weights (GELU), zero biases; SplitMix64 draws
        std::vector<float> h(nparam, 0.0f);
        uint64_t s = seed;
        auto next = [&s]() {
            uint64_t z = (s += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return static_cast<double>(z >> 11) * 0x1.0p-53;
        };
        for (int l = 0; l < 3; l++) {
            const double lim = std::sqrt(6.0 / d[l]);
            for (size_t e = 0; e < static_cast<size_t>(d[l]) * d[l + 1]; e++) {
                h[woff[l] + e] = static_cast<float>((2.0 * next() - 1.0) * lim);
            }
        }
        set_params(h.data());
        std::vector<float> ones(d3, 1.0f);
        HIP_OK(hipMemcpy(colw, ones.data(), d3 * sizeof(float), hipMemcpyHostToDevice));
    }

    ~Mlp() {
        for (void *p : {static_cast<void *>(p32), static_cast<void *>(grad), static_cast<void *>(mom),
                        static_cast<void *>(vel), static_cast<void *>(pbf), static_cast<void *>(xf),
                        static_cast<void *>(xb), static_cast<void *>(a1), static_cast<void *>(a2),
                        static_cast<void *>(z1), static_cast<void *>(z2), static_cast<void *>(dz1),
                        static_cast<void *>(dz2), static_cast<void *>(y), static_cast<void *>(dy),
                        static_cast<void *>(tb), static_cast<void *>(colw), static_cast<void *>(lpart),
                        static_cast<void *>(data_x), static_cast<void *>(data_t)}) {
            if (p) {
                (void)hipFree(p);
            }
        }
        if (stream) {
            (void)hipStreamDestroy(stream);
        }
    }

    static unsigned blocks(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

    void set_params(const float *host) {
        HIP_OK(hipMemcpyAsync(p32, host, nparam * sizeof(float), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_to_bf16, dim3(blocks(nparam)), dim3(256), 0, stream, p32, pbf,
                           static_cast<int64_t>(nparam));
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemsetAsync(mom, 0, nparam * sizeof(float), stream));
        HIP_OK(hipMemsetAsync(vel, 0, nparam * sizeof(float), stream));
        HIP_OK(hipStreamSynchronize(stream));
        step = 0;
    }

    template <class TA, class TB, bool TRA, bool TRB> void gemm(const GemmArgs &g) {
        const dim3 grid((g.N + 63) / 64, (g.M + 63) / 64);
        hipLaunchKernelGGL((k_gemm_bf16<TA, TB, TRA, TRB>), grid, dim3(256), 0, stream, g);
        HIP_OK(hipGetLastError());
    }

    // y[B x d3] for BF16 features x[B x d0]
    void forward(const uint16_t *x, int B) {
        GemmArgs g{};
        g = GemmArgs{x, d[0], pbf + woff[0], d[1], nullptr, d[1], B, d[1], d[0], p32 + boff[0], nullptr, z1,
                     a1, d[1], kEpiGelu};
        gemm<uint16_t, uint16_t, false, false>(g);
        g = GemmArgs{a1, d[1], pbf + woff[1], d[2], nullptr, d[2], B, d[2], d[1], p32 + boff[1], nullptr, z2,
                     a2, d[2], kEpiGelu};
        gemm<uint16_t, uint16_t, false, false>(g);
        g = GemmArgs{a2, d[2], pbf + woff[2], d[3], y, d[3], B, d[3], d[2], p32 + boff[2], nullptr, nullptr,
                     nullptr, 0, kEpiStore};
        gemm<uint16_t, uint16_t, false, false>(g);
    }

    // gradients of the weighted MSE against targets t[B x d3] into `grad`;
    // returns the loss (requires forward(x, B) first)
    double backward(const uint16_t *x, const float *t, int B) {
        hipLaunchKernelGGL(k_mse, dim3(kLossBlocks), dim3(256), 0, stream, y, t, B, d[3], colw, dy, lpart);
        HIP_OK(hipGetLastError());
        // layer 3
        gemm<uint16_t, float, true, false>(GemmArgs{a2, d[2], dy, d[3], grad + woff[2], d[3], d[2], d[3], B,
                                                    nullptr, nullptr, nullptr, nullptr, 0, kEpiStore});
        hipLaunchKernelGGL(k_colsum, dim3(blocks(d[3])), dim3(256), 0, stream, dy, B, d[3], d[3], grad + boff[2]);
        gemm<float, uint16_t, false, true>(GemmArgs{dy, d[3], pbf + woff[2], d[3], dz2, d[2], B, d[2], d[3],
                                                    nullptr, z2, nullptr, nullptr, 0, kEpiGeluGrad});
        // layer 2
        gemm<uint16_t, float, true, false>(GemmArgs{a1, d[1], dz2, d[2], grad + woff[1], d[2], d[1], d[2], B,
                                                    nullptr, nullptr, nullptr, nullptr, 0, kEpiStore});
        hipLaunchKernelGGL(k_colsum, dim3(blocks(d[2])), dim3(256), 0, stream, dz2, B, d[2], d[2], grad + boff[1]);
        gemm<float, uint16_t, false, true>(GemmArgs{dz2, d[2], pbf + woff[1], d[2], dz1, d[1], B, d[1], d[2],
                                                    nullptr, z1, nullptr, nullptr, 0, kEpiGeluGrad});
        // layer 1
        gemm<uint16_t, float, true, false>(GemmArgs{x, d[0], dz1, d[1], grad + woff[0], d[1], d[0], d[1], B,
                                                    nullptr, nullptr, nullptr, nullptr, 0, kEpiStore});
        hipLaunchKernelGGL(k_colsum, dim3(blocks(d[1])), dim3(256), 0, stream, dz1, B, d[1], d[1], grad + boff[0]);
        HIP_OK(hipGetLastError());
        std::vector<double> part(kLossBlocks);
        HIP_OK(hipMemcpyAsync(part.data(), lpart, kLossBlocks * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        double s = 0;
        for (double v : part) {
            s += v;
        }
        return s / B;
    }

    void adam(float lr) {
        step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, static_cast<float>(step));
        const float c2 = 1.0f - std::pow(b2, static_cast<float>(step));
        hipLaunchKernelGGL(k_adam, dim3(blocks(nparam)), dim3(256), 0, stream, p32, pbf, mom, vel, grad,
                           static_cast<int64_t>(nparam), lr, b1, b2, 1e-8f, c1, c2);
        HIP_OK(hipGetLastError());
    }

    void check_batch(int64_t n) const {
        if (n < 1 || n > maxb) {
            throw InvalidArg("batch size must be in [1, " + std::to_string(maxb) + "]");
        }
    }

    void upload_x(const float *x_host, int n) {
        check_batch(n);
        HIP_OK(hipMemcpyAsync(xf, x_host, static_cast<size_t>(n) * d[0] * sizeof(float), hipMemcpyHostToDevice,
                              stream));
        hipLaunchKernelGGL(k_to_bf16, dim3(blocks(static_cast<int64_t>(n) * d[0])), dim3(256), 0, stream, xf, xb,
                           static_cast<int64_t>(n) * d[0]);
        HIP_OK(hipGetLastError());
    }

    void predict(const float *x_host, int n, float *y_host) {
        upload_x(x_host, n);
        forward(xb, n);
        HIP_OK(hipMemcpyAsync(y_host, y, static_cast<size_t>(n) * d[3] * sizeof(float), hipMemcpyDeviceToHost,
                              stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    double gradients(const float *x_host, const float *t_host, int n, float *g_host) {
        upload_x(x_host, n);
        HIP_OK(hipMemcpyAsync(tb, t_host, static_cast<size_t>(n) * d[3] * sizeof(float), hipMemcpyHostToDevice,
                              stream));
        forward(xb, n);
        const double loss = backward(xb, tb, n);
        HIP_OK(hipMemcpy(g_host, grad, nparam * sizeof(float), hipMemcpyDeviceToHost));
        return loss;
    }

    void load_data(const float *x_host, const float *t_host, int64_t n) {
        for (void *p : {static_cast<void *>(data_x), static_cast<void *>(data_t)}) {
            if (p) {
                HIP_OK(hipFree(p));
            }
        }
        data_x = nullptr;
        data_t = nullptr;
        ndata = n;
        cursor = 0;
        HIP_OK(hipMalloc(reinterpret_cast<void **>(&data_x), std::max<int64_t>(n, 1) * d[0] * sizeof(uint16_t)));
        HIP_OK(hipMalloc(reinterpret_cast<void **>(&data_t), std::max<int64_t>(n, 1) * d[3] * sizeof(float)));
        // features through the FP32 staging buffer, maxb rows at a time
        for (int64_t r = 0; r < n; r += maxb) {
            const int64_t m = std::min<int64_t>(maxb, n - r);
            HIP_OK(hipMemcpyAsync(xf, x_host + r * d[0], m * d[0] * sizeof(float), hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL(k_to_bf16, dim3(blocks(m * d[0])), dim3(256), 0, stream, xf, data_x + r * d[0],
                               m * d[0]);
            HIP_OK(hipGetLastError());
            HIP_OK(hipStreamSynchronize(stream));
        }
        HIP_OK(hipMemcpy(data_t, t_host, n * d[3] * sizeof(float), hipMemcpyHostToDevice));
    }

    // `steps` Adam steps over consecutive minibatches of the loaded data
    double train(int64_t steps, int batch, float lr) {
        check_batch(batch);
        if (ndata < batch) {
            throw InvalidArg("load at least one batch of data first");
        }
        double tot = 0;
        for (int64_t s = 0; s < steps; s++) {
            if (cursor + batch > ndata) {
                cursor = 0;
            }
            const uint16_t *x = data_x + cursor * d[0];
            const float *t = data_t + cursor * d[3];
            cursor += batch;
            forward(x, batch);
            tot += backward(x, t, batch);
            adam(lr);
        }
        HIP_OK(hipStreamSynchronize(stream));
        return steps ? tot / steps : 0.0;
    }
};

} // namespace plamd
