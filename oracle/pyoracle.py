"""ctypes front end for the two CPU oracles (test infrastructure only).

``Oracle()`` loads this repository's restatement (oracle/liblq_oracle.so);
``Oracle(ref=True)`` loads the reference's own Lightning-Qubit code compiled
in place (oracle/_ref/libref_lq.so).  Both export the interface declared in
oracle/oracle_api.h.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.
"""
import ctypes
import os

import numpy as np

from lightning_unification_amd.circuits import encode_observables
from lightning_unification_amd.gates import GATE_ID

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liblq_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_lq.so")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleError(RuntimeError):
    pass


class Oracle:
    def __init__(self, ref=False):
        path = REF_SO if ref else ORACLE_SO
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.lib.orc_last_error.restype = ctypes.c_char_p
        self.ref = ref

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(self.lib.orc_last_error().decode())

    @staticmethod
    def _suf(sv):
        if sv.dtype == np.complex128:
            return "c128", np.float64
        if sv.dtype == np.complex64:
            return "c64", np.float32
        raise TypeError(sv.dtype)

    @staticmethod
    def nqubits(sv):
        n = int(sv.size).bit_length() - 1
        assert sv.size == 1 << n
        return n

    def num_threads(self):
        return self.lib.orc_num_threads()

    def apply_tape(self, sv, tape):
        suf, _ = self._suf(sv)
        ids, nw, wires, inv, npar, par = tape.encode()
        wires = wires if wires.size else np.zeros(1, np.int32)
        par = par if par.size else np.zeros(1)
        fn = getattr(self.lib, f"orc_apply_ops_{suf}")
        self._chk(fn(_p(sv), self.nqubits(sv), len(ids), _p(ids), _p(nw), _p(wires), _p(inv),
                     _p(npar), _p(par)))
        return sv

    def apply_matrix(self, sv, mat, wires, inverse=False):
        suf, _ = self._suf(sv)
        m = np.ascontiguousarray(mat, sv.dtype)
        w = np.array(wires, np.int32)
        fn = getattr(self.lib, f"orc_apply_matrix_{suf}")
        self._chk(fn(_p(sv), self.nqubits(sv), _p(m), len(w), _p(w), int(inverse)))
        return sv

    def apply_generator(self, sv, name, wires, adj=False):
        suf, _ = self._suf(sv)
        w = np.array(wires, np.int32)
        scale = ctypes.c_double()
        fn = getattr(self.lib, f"orc_apply_generator_{suf}")
        self._chk(fn(_p(sv), self.nqubits(sv), GATE_ID[name], len(w), _p(w), int(adj),
                     ctypes.byref(scale)))
        return scale.value

    def _obs_args(self, sv, obs_list):
        prog, coeffs, mats = encode_observables(obs_list)
        return prog, coeffs, np.ascontiguousarray(mats.astype(sv.dtype))

    def expval(self, sv, ob, var=False):
        suf, _ = self._suf(sv)
        prog, coeffs, mats = self._obs_args(sv, [ob])
        out = ctypes.c_double()
        fn = getattr(self.lib, f"orc_{'var' if var else 'expval'}_{suf}")
        self._chk(fn(_p(sv), self.nqubits(sv), _p(prog), _p(coeffs), _p(mats), ctypes.byref(out)))
        return out.value

    def var(self, sv, ob):
        return self.expval(sv, ob, var=True)

    def expval_named(self, sv, name, wires, var=False):
        suf, _ = self._suf(sv)
        w = np.array(wires, np.int32)
        out = ctypes.c_double()
        fn = getattr(self.lib, f"orc_{'var' if var else 'expval'}_named_{suf}")
        self._chk(fn(_p(sv), self.nqubits(sv), GATE_ID[name], len(w), _p(w), ctypes.byref(out)))
        return out.value

    def expval_matrix(self, sv, mat, wires):
        suf, _ = self._suf(sv)
        m = np.ascontiguousarray(mat, sv.dtype)
        w = np.array(wires, np.int32)
        out = ctypes.c_double()
        fn = getattr(self.lib, f"orc_expval_matrix_{suf}")
        self._chk(fn(_p(sv), self.nqubits(sv), _p(m), len(w), _p(w), ctypes.byref(out)))
        return out.value

    def probs(self, sv, wires=None):
        suf, real = self._suf(sv)
        n = self.nqubits(sv)
        if wires is None:
            out = np.zeros(sv.size, real)
            w = np.zeros(1, np.int32)
            nw = -1
        else:
            w = np.array(wires if len(wires) else [0], np.int32)
            nw = len(wires)
            out = np.zeros(1 << nw, real)
        fn = getattr(self.lib, f"orc_probs_{suf}")
        self._chk(fn(_p(sv), n, nw, _p(w), _p(out)))
        return out

    def sample_cdf(self, sv, uniforms):
        suf, _ = self._suf(sv)
        u = np.ascontiguousarray(uniforms, np.float64)
        idx = np.zeros(u.size, np.int64)
        fn = getattr(self.lib, f"orc_sample_cdf_{suf}")
        self._chk(fn(_p(sv), self.nqubits(sv), ctypes.c_int64(u.size), _p(u), _p(idx)))
        return idx

    def adjoint(self, sv, tape, obs_list, tp):
        suf, real = self._suf(sv)
        ids, nw, wires, inv, npar, par = tape.encode()
        wires = wires if wires.size else np.zeros(1, np.int32)
        par = par if par.size else np.zeros(1)
        prog, coeffs, mats = self._obs_args(sv, obs_list)
        tpa = np.array(tp if len(tp) else [0], np.int32)
        jac = np.zeros(max(1, len(obs_list) * len(tp)), real)
        fn = getattr(self.lib, f"orc_adjoint_{suf}")
        self._chk(fn(_p(sv), self.nqubits(sv), len(ids), _p(ids), _p(nw), _p(wires), _p(inv),
                     _p(npar), _p(par), _p(prog), _p(coeffs), _p(mats), len(tp), _p(tpa),
                     _p(jac)))
        return jac[: len(obs_list) * len(tp)].reshape(len(obs_list), len(tp))

    def vjp(self, sv, tape, dy, tp):
        """<dy| d psi / d theta_t> (VectorJacobianProduct.hpp:77-155)."""
        suf, real = self._suf(sv)
        ids, nw, wires, inv, npar, par = tape.encode()
        wires = wires if wires.size else np.zeros(1, np.int32)
        par = par if par.size else np.zeros(1)
        d = np.ascontiguousarray(np.asarray(dy, sv.dtype))
        tpa = np.array(tp if len(tp) else [0], np.int32)
        out = np.zeros(2 * max(1, len(tp)), real)
        fn = getattr(self.lib, f"orc_vjp_{suf}")
        self._chk(fn(_p(sv), self.nqubits(sv), len(ids), _p(ids), _p(nw), _p(wires), _p(inv),
                     _p(npar), _p(par), _p(d), len(tp), _p(tpa), _p(out)))
        return out[:2 * len(tp)].astype(np.float64).view(np.complex128)

    def sparse(self, sv, row_map, entries, values, var=False):
        """<psi|H|psi> or Var(H) for a CSR matrix (MeasurementsLQubit.hpp:183-214,437-471)."""
        suf, _ = self._suf(sv)
        rm = np.ascontiguousarray(row_map, np.int64)
        en = np.ascontiguousarray(entries if len(entries) else [0], np.int64)
        va = np.ascontiguousarray(values if len(values) else [0], sv.dtype)
        out = ctypes.c_double()
        fn = getattr(self.lib, f"orc_{'var' if var else 'expval'}_sparse_{suf}")
        self._chk(fn(_p(sv), self.nqubits(sv), _p(rm), ctypes.c_int64(rm.size), _p(en), _p(va),
                     ctypes.c_int64(len(values)), ctypes.byref(out)))
        return out.value

