// C-ABI of the batched layered-circuit engine (pl_batch_*, include/pl_amd.h):
// many independent circuits, forward pass, energy and adjoint gradients in
// one replayed hipGraph (batch_host.hpp, kernels_batch.hpp).
#include "engine_common.hpp"
#include "batch_host.hpp"

using namespace plamd;

struct pl_batch {
    std::unique_ptr<BatchPlanBase> plan;
};

int pl_batch_create(int dtype, int ncirc, const int *nqubits, const int *depth, const char *pattern,
                    int entangler, const int *prog, const double *coeffs, int tile_width,
                    pl_batch_t *out) {
    return guarded([&] {
        if (!out || !nqubits || !depth || !pattern || !prog) {
            throw InvalidArg("null argument");
        }
        if (ncirc < 1) {
            throw InvalidArg("need at least one circuit");
        }
        if (dtype != PL_C64 && dtype != PL_C128) {
            throw InvalidArg("dtype must be PL_C64 or PL_C128");
        }
        HIP_OK(hipSetDevice(0));
        ObsReader rd{prog, coeffs, nullptr};
        if (*rd.prog++ != ncirc) {
            throw InvalidArg("need one observable per circuit");
        }
        std::vector<std::vector<PauliTerm>> terms(ncirc);
        std::vector<std::vector<uint64_t>> xm(ncirc);
        for (int j = 0; j < ncirc; j++) {
            const ObsNode o = rd.one();
            std::vector<const ObsNode *> parts;
            std::vector<double> cs;
            if (o.kind == 3) {
                for (size_t t = 0; t < o.kids.size(); t++) {
                    parts.push_back(&o.kids[t]);
                    cs.push_back(o.coeffs[t]);
                }
            } else {
                parts.push_back(&o);
                cs.push_back(1.0);
            }
            for (size_t t = 0; t < parts.size(); t++) {
                PauliWord pw;
                if (!as_pauli(*parts[t], nqubits[j], pw)) {
                    throw InvalidArg("batched observables must be sums of Pauli words");
                }
                terms[j].push_back(PauliTerm{pw.zy, cs[t], pw.ny & 3, 0});
                xm[j].push_back(pw.x);
            }
        }
        auto b = std::make_unique<pl_batch>();
        if (dtype == PL_C128) {
            b->plan = std::make_unique<BatchPlan<double>>(ncirc, nqubits, depth, pattern, entangler,
                                                          terms, xm, tile_width);
        } else {
            b->plan = std::make_unique<BatchPlan<float>>(ncirc, nqubits, depth, pattern, entangler,
                                                         terms, xm, tile_width);
        }
        *out = b.release();
    });
}

int pl_batch_run(pl_batch_t b, const double *params, double *energies, double *grads) {
    return guarded([&] {
        if (!b || !energies || (!params && b->plan->total_params() > 0)) {
            throw InvalidArg("null argument");
        }
        HIP_OK(hipSetDevice(0));
        b->plan->run(params, energies, grads);
    });
}

int pl_batch_info(pl_batch_t b, int *tile_width, int *launches, int64_t *nparams) {
    return guarded([&] {
        if (!b) throw InvalidArg("null argument");
        if (tile_width) *tile_width = b->plan->tile_width();
        if (launches) *launches = b->plan->launches();
        if (nparams) *nparams = static_cast<int64_t>(b->plan->total_params());
    });
}

int pl_batch_destroy(pl_batch_t b) {
    return guarded([&] { delete b; });
}

