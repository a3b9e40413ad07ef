/*
 * Driver over the reference Lightning-Qubit C++ sources, compiled in place
 * from /root/reference by oracle/Makefile into oracle/_ref/libref_lq.so.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle_api.h).  This file is this
 * repository's own code: it only calls the reference's public C++ API the way
 * its pybind11 layer does (core/src/bindings/Bindings.hpp:100-116, 405-476),
 * over StateVectorLQubitRaw, the backend the bindings register
 * (simulators/lightning_qubit/bindings/LQubitBindings.hpp:45-47).
 */
#include <complex>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "AdjointJacobianLQubit.hpp"
#include "Constant.hpp"
#include "ConstantUtil.hpp"
#include "GateOperation.hpp"
#include "JacobianData.hpp"
#include "MeasurementsLQubit.hpp"
#include "ObservablesLQubit.hpp"
#include "StateVectorLQubitRaw.hpp"
#include "VectorJacobianProduct.hpp"

#include "oracle_api.h"

using Pennylane::Gates::GateOperation;
using Pennylane::LightningQubit::StateVectorLQubitRaw;
namespace LQO = Pennylane::LightningQubit::Observables;
namespace LQM = Pennylane::LightningQubit::Measures;
namespace LQA = Pennylane::LightningQubit::Algorithms;

namespace {
thread_local std::string g_err;

std::string gate_name(int id) {
    if (id == -1) {
        return "QubitStateVector";
    }
    return std::string(Pennylane::Util::lookup(
        Pennylane::Gates::Constant::gate_names, static_cast<GateOperation>(id)));
}

template <class F> int guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
    } catch (...) {
        g_err = "unknown error";
    }
    return -1;
}

template <class P> using SV = StateVectorLQubitRaw<P>;
template <class P> using ObsPtr = std::shared_ptr<Pennylane::Observables::Observable<SV<P>>>;

/* Parse the observable program of oracle_api.h into reference objects. */
template <class P> struct ObsParser {
    const int *prog;
    const double *coeffs;
    const P *mats;

    ObsPtr<P> factor() {
        const int kind = *prog++;
        const int op = *prog++;
        const int nw = *prog++;
        std::vector<size_t> w(prog, prog + nw);
        prog += nw;
        if (kind == 0) {
            return std::make_shared<LQO::NamedObs<SV<P>>>(gate_name(op), w);
        }
        const size_t dim = size_t{1} << (2 * nw);
        std::vector<std::complex<P>> m(dim);
        for (size_t i = 0; i < dim; i++) {
            m[i] = {mats[2 * i], mats[2 * i + 1]};
        }
        mats += 2 * dim;
        return std::make_shared<LQO::HermitianObs<SV<P>>>(m, w);
    }
    ObsPtr<P> term() {
        const int nf = *prog++;
        if (nf == 1) {
            return factor();
        }
        std::vector<ObsPtr<P>> fs;
        for (int f = 0; f < nf; f++) {
            fs.push_back(factor());
        }
        return LQO::TensorProdObs<SV<P>>::create(fs);
    }
    ObsPtr<P> one() {
        const int is_ham = *prog++;
        const int nterms = *prog++;
        if (!is_ham) {
            if (nterms != 1) {
                throw std::invalid_argument("non-Hamiltonian observable must have one term");
            }
            return term();
        }
        std::vector<P> c;
        std::vector<ObsPtr<P>> ts;
        for (int t = 0; t < nterms; t++) {
            c.push_back(static_cast<P>(*coeffs++));
            ts.push_back(term());
        }
        return std::make_shared<LQO::Hamiltonian<SV<P>>>(c, ts);
    }
    std::vector<ObsPtr<P>> all() {
        const int nobs = *prog++;
        std::vector<ObsPtr<P>> v;
        for (int i = 0; i < nobs; i++) {
            v.push_back(one());
        }
        return v;
    }
};

template <class P>
Pennylane::Algorithms::OpsData<SV<P>>
make_ops(int nops, const int *ids, const int *nwires, const int *wires,
         const uint8_t *inverse, const int *nparams, const double *params) {
    std::vector<std::string> names;
    std::vector<std::vector<P>> ps;
    std::vector<std::vector<size_t>> ws;
    std::vector<bool> inv;
    for (int i = 0; i < nops; i++) {
        names.push_back(gate_name(ids[i]));
        ws.emplace_back(wires, wires + nwires[i]);
        wires += nwires[i];
        std::vector<P> p;
        for (int j = 0; j < nparams[i]; j++) {
            p.push_back(static_cast<P>(params[j]));
        }
        params += nparams[i];
        ps.push_back(p);
        inv.push_back(inverse[i] != 0);
    }
    return {names, ps, ws, inv};
}

template <class P> std::complex<P> *as_c(P *p) {
    return reinterpret_cast<std::complex<P> *>(p);
}

template <class P>
int apply_ops(P *sv, int n, int nops, const int *ids, const int *nwires,
              const int *wires, const uint8_t *inverse, const int *nparams,
              const double *params) {
    return guard([&] {
        SV<P> s(as_c(sv), size_t{1} << n);
        for (int i = 0; i < nops; i++) {
            std::vector<size_t> w(wires, wires + nwires[i]);
            wires += nwires[i];
            std::vector<P> p;
            for (int j = 0; j < nparams[i]; j++) {
                p.push_back(static_cast<P>(params[j]));
            }
            params += nparams[i];
            s.applyOperation(gate_name(ids[i]), w, inverse[i] != 0, p);
        }
    });
}

template <class P>
int apply_matrix(P *sv, int n, const P *mat, int nw, const int *wires, int inverse) {
    return guard([&] {
        SV<P> s(as_c(sv), size_t{1} << n);
        std::vector<size_t> w(wires, wires + nw);
        s.applyMatrix(reinterpret_cast<const std::complex<P> *>(mat), w, inverse != 0);
    });
}

template <class P>
int apply_generator(P *sv, int n, int id, int nw, const int *wires, int adj, double *scale) {
    return guard([&] {
        SV<P> s(as_c(sv), size_t{1} << n);
        std::vector<size_t> w(wires, wires + nw);
        *scale = s.applyGenerator(gate_name(id), w, adj != 0);
    });
}

template <class P>
int expval_obs(const P *sv, int n, const int *prog, const double *coeffs,
               const P *mats, double *out, bool var) {
    return guard([&] {
        SV<P> s(as_c(const_cast<P *>(sv)), size_t{1} << n);
        ObsParser<P> parser{prog + 1, coeffs, mats}; // single observable
        auto ob = parser.one();
        LQM::Measurements<SV<P>> m(s);
        *out = var ? m.var(*ob) : m.expval(*ob);
    });
}

template <class P>
int named(const P *sv, int n, int id, int nw, const int *wires, double *out, bool var) {
    return guard([&] {
        SV<P> s(as_c(const_cast<P *>(sv)), size_t{1} << n);
        LQM::Measurements<SV<P>> m(s);
        std::vector<size_t> w(wires, wires + nw);
        *out = var ? m.var(gate_name(id), w) : m.expval(gate_name(id), w);
    });
}

template <class P>
int expval_matrix(const P *sv, int n, const P *mat, int nw, const int *wires, double *out) {
    return guard([&] {
        SV<P> s(as_c(const_cast<P *>(sv)), size_t{1} << n);
        LQM::Measurements<SV<P>> m(s);
        std::vector<size_t> w(wires, wires + nw);
        const size_t dim = size_t{1} << (2 * nw);
        std::vector<std::complex<P>> mm(reinterpret_cast<const std::complex<P> *>(mat),
                                        reinterpret_cast<const std::complex<P> *>(mat) + dim);
        *out = m.expval(mm, w);
    });
}

template <class P>
int probs(const P *sv, int n, int nw, const int *wires, P *out) {
    return guard([&] {
        SV<P> s(as_c(const_cast<P *>(sv)), size_t{1} << n);
        LQM::Measurements<SV<P>> m(s);
        std::vector<P> r;
        if (nw < 0) {
            r = m.probs();
        } else {
            std::vector<size_t> w(wires, wires + nw);
            r = m.probs(w);
        }
        std::copy(r.begin(), r.end(), out);
    });
}

template <class P>
int adjoint(const P *sv, int n, int nops, const int *ids, const int *nwires,
            const int *wires, const uint8_t *inverse, const int *nparams,
            const double *params, const int *prog, const double *coeffs,
            const P *mats, int ntp, const int *tp, P *jac) {
    return guard([&] {
        SV<P> s(as_c(const_cast<P *>(sv)), size_t{1} << n);
        ObsParser<P> parser{prog, coeffs, mats};
        auto obs = parser.all();
        auto ops = make_ops<P>(nops, ids, nwires, wires, inverse, nparams, params);
        std::vector<size_t> tps(tp, tp + ntp);
        std::vector<P> j(obs.size() * tps.size(), P{0});
        Pennylane::Algorithms::JacobianData<SV<P>> jd{
            ops.getTotalNumParams(), s.getLength(), s.getData(), obs, ops, tps};
        LQA::AdjointJacobian<SV<P>> adj;
        adj.adjointJacobian(std::span{j}, jd);
        std::copy(j.begin(), j.end(), jac);
    });
}
template <class P>
int vjp(const P *sv, int n, int nops, const int *ids, const int *nwires, const int *wires,
        const uint8_t *inverse, const int *nparams, const double *params, const P *dy, int ntp,
        const int *tp, P *out) {
    return guard([&] {
        SV<P> s(as_c(const_cast<P *>(sv)), size_t{1} << n);
        auto ops = make_ops<P>(nops, ids, nwires, wires, inverse, nparams, params);
        std::vector<size_t> tps(tp, tp + ntp);
        std::vector<std::complex<P>> j(tps.size());
        Pennylane::Algorithms::JacobianData<SV<P>> jd{ops.getTotalNumParams(), s.getLength(),
                                                      s.getData(), {}, ops, tps};
        LQA::VectorJacobianProduct<SV<P>> v;
        const auto *d = reinterpret_cast<const std::complex<P> *>(dy);
        v(std::span{j}, jd, std::span<const std::complex<P>>{d, s.getLength()});
        for (size_t t = 0; t < j.size(); t++) {
            out[2 * t] = j[t].real();
            out[2 * t + 1] = j[t].imag();
        }
    });
}

template <class P>
int sparse(const P *sv, int n, const int64_t *rm, int64_t rms, const int64_t *en, const P *va,
           int64_t nnz, double *out, bool var) {
    return guard([&] {
        SV<P> s(as_c(const_cast<P *>(sv)), size_t{1} << n);
        Pennylane::LightningQubit::Measures::Measurements<SV<P>> m(s);
        const auto *v = reinterpret_cast<const std::complex<P> *>(va);
        *out = var ? m.var(rm, rms, en, v, nnz) : m.expval(rm, rms, en, v, nnz);
    });
}
} // namespace

extern "C" {
const char *orc_last_error(void) { return g_err.c_str(); }
int orc_num_threads(void) { return 0; }

#define ORC_DEF(SUF, REAL)                                                              \
    int orc_apply_ops_##SUF(REAL *sv, int n, int nops, const int *ids,                  \
                            const int *nwires, const int *wires,                        \
                            const uint8_t *inverse, const int *nparams,                 \
                            const double *params) {                                     \
        return apply_ops<REAL>(sv, n, nops, ids, nwires, wires, inverse, nparams,       \
                               params);                                                 \
    }                                                                                   \
    int orc_apply_matrix_##SUF(REAL *sv, int n, const REAL *mat, int nw,                \
                               const int *wires, int inverse) {                         \
        return apply_matrix<REAL>(sv, n, mat, nw, wires, inverse);                      \
    }                                                                                   \
    int orc_apply_generator_##SUF(REAL *sv, int n, int id, int nw, const int *wires,    \
                                  int adj, double *scale) {                             \
        return apply_generator<REAL>(sv, n, id, nw, wires, adj, scale);                 \
    }                                                                                   \
    int orc_expval_##SUF(const REAL *sv, int n, const int *prog, const double *coeffs,  \
                         const REAL *mats, double *out) {                               \
        return expval_obs<REAL>(sv, n, prog, coeffs, mats, out, false);                 \
    }                                                                                   \
    int orc_var_##SUF(const REAL *sv, int n, const int *prog, const double *coeffs,     \
                      const REAL *mats, double *out) {                                  \
        return expval_obs<REAL>(sv, n, prog, coeffs, mats, out, true);                  \
    }                                                                                   \
    int orc_expval_named_##SUF(const REAL *sv, int n, int id, int nw, const int *wires, \
                               double *out) {                                           \
        return named<REAL>(sv, n, id, nw, wires, out, false);                           \
    }                                                                                   \
    int orc_var_named_##SUF(const REAL *sv, int n, int id, int nw, const int *wires,    \
                            double *out) {                                              \
        return named<REAL>(sv, n, id, nw, wires, out, true);                            \
    }                                                                                   \
    int orc_expval_matrix_##SUF(const REAL *sv, int n, const REAL *mat, int nw,         \
                                const int *wires, double *out) {                        \
        return expval_matrix<REAL>(sv, n, mat, nw, wires, out);                         \
    }                                                                                   \
    int orc_probs_##SUF(const REAL *sv, int n, int nw, const int *wires, REAL *out) {   \
        return probs<REAL>(sv, n, nw, wires, out);                                      \
    }                                                                                   \
    int orc_adjoint_##SUF(const REAL *sv, int n, int nops, const int *ids,              \
                          const int *nwires, const int *wires, const uint8_t *inverse,  \
                          const int *nparams, const double *params, const int *prog,    \
                          const double *coeffs, const REAL *mats, int ntp,              \
                          const int *tp, REAL *jac) {                                   \
        return adjoint<REAL>(sv, n, nops, ids, nwires, wires, inverse, nparams, params, \
                             prog, coeffs, mats, ntp, tp, jac);                         \
    }

ORC_DEF(c128, double)
ORC_DEF(c64, float)

int orc_vjp_c128(const double *sv, int n, int nops, const int *ids, const int *nw, const int *w,
                 const uint8_t *iv, const int *np, const double *pr, const double *dy, int ntp,
                 const int *tp, double *out) {
    return vjp<double>(sv, n, nops, ids, nw, w, iv, np, pr, dy, ntp, tp, out);
}
int orc_vjp_c64(const float *sv, int n, int nops, const int *ids, const int *nw, const int *w,
                const uint8_t *iv, const int *np, const double *pr, const float *dy, int ntp,
                const int *tp, float *out) {
    return vjp<float>(sv, n, nops, ids, nw, w, iv, np, pr, dy, ntp, tp, out);
}
int orc_expval_sparse_c128(const double *sv, int n, const int64_t *rm, int64_t rms,
                           const int64_t *en, const double *va, int64_t nnz, double *out) {
    return sparse<double>(sv, n, rm, rms, en, va, nnz, out, false);
}
int orc_expval_sparse_c64(const float *sv, int n, const int64_t *rm, int64_t rms,
                          const int64_t *en, const float *va, int64_t nnz, double *out) {
    return sparse<float>(sv, n, rm, rms, en, va, nnz, out, false);
}
int orc_var_sparse_c128(const double *sv, int n, const int64_t *rm, int64_t rms,
                        const int64_t *en, const double *va, int64_t nnz, double *out) {
    return sparse<double>(sv, n, rm, rms, en, va, nnz, out, true);
}
int orc_var_sparse_c64(const float *sv, int n, const int64_t *rm, int64_t rms,
                       const int64_t *en, const float *va, int64_t nnz, double *out) {
    return sparse<float>(sv, n, rm, rms, en, va, nnz, out, true);
}
}
