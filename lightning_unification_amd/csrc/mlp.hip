// C-ABI of the surrogate MLP (pl_mlp_*, include/pl_amd.h): BF16 matrix-core
// forward and backward passes, FP32 master weights and Adam (mlp_host.hpp,
// kernels_mlp.hpp).
#include "engine_common.hpp"
#include "mlp_host.hpp"

using namespace plamd;

struct pl_mlp {
    std::unique_ptr<plamd::Mlp> net;
};

int pl_mlp_create(int d0, int d1, int d2, int d3, int max_batch, uint64_t seed, pl_mlp_t *out) {
    return guarded([&] {
        if (!out || d0 < 1 || d1 < 1 || d2 < 1 || d3 < 1 || max_batch < 1 || max_batch > (1 << 20)) {
            throw InvalidArg("invalid MLP shape");
        }
        HIP_OK(hipSetDevice(0));
        auto m = std::make_unique<pl_mlp>();
        m->net = std::make_unique<Mlp>(d0, d1, d2, d3, max_batch, seed);
        *out = m.release();
    });
}

int pl_mlp_destroy(pl_mlp_t m) {
    return guarded([&] { delete m; });
}

static Mlp &mlp_of(pl_mlp_t m) {
    if (!m) {
        throw InvalidArg("null MLP handle");
    }
    HIP_OK(hipSetDevice(0));
    return *m->net;
}

int pl_mlp_num_params(pl_mlp_t m, int64_t *out) {
    return guarded([&] {
        if (!out) throw InvalidArg("null output");
        *out = static_cast<int64_t>(mlp_of(m).nparam);
    });
}

int pl_mlp_get_params(pl_mlp_t m, float *params) {
    return guarded([&] {
        Mlp &net = mlp_of(m);
        if (!params) throw InvalidArg("null output");
        HIP_OK(hipStreamSynchronize(net.stream));
        HIP_OK(hipMemcpy(params, net.p32, net.nparam * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int pl_mlp_set_params(pl_mlp_t m, const float *params) {
    return guarded([&] {
        if (!params) throw InvalidArg("null input");
        mlp_of(m).set_params(params);
    });
}

int pl_mlp_set_loss_weights(pl_mlp_t m, const float *w) {
    return guarded([&] {
        Mlp &net = mlp_of(m);
        if (!w) throw InvalidArg("null input");
        HIP_OK(hipStreamSynchronize(net.stream));
        HIP_OK(hipMemcpy(net.colw, w, net.d[3] * sizeof(float), hipMemcpyHostToDevice));
    });
}

int pl_mlp_predict(pl_mlp_t m, const float *x, int n, float *y) {
    return guarded([&] {
        if (!x || !y) throw InvalidArg("null argument");
        mlp_of(m).predict(x, n, y);
    });
}

int pl_mlp_gradients(pl_mlp_t m, const float *x, const float *t, int n, double *loss, float *grads) {
    return guarded([&] {
        if (!x || !t || !grads) throw InvalidArg("null argument");
        const double l = mlp_of(m).gradients(x, t, n, grads);
        if (loss) *loss = l;
    });
}

int pl_mlp_load_data(pl_mlp_t m, const float *x, const float *t, int64_t n) {
    return guarded([&] {
        if (!x || !t || n < 1) throw InvalidArg("invalid training set");
        mlp_of(m).load_data(x, t, n);
    });
}

int pl_mlp_train(pl_mlp_t m, int64_t steps, int batch, float lr, double *mean_loss) {
    return guarded([&] {
        if (steps < 0) throw InvalidArg("negative step count");
        const double l = mlp_of(m).train(steps, batch, lr);
        if (mean_loss) *mean_loss = l;
    });
}

