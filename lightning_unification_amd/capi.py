"""ctypes binding of the C-ABI in include/pl_amd.h (libpl_amd.so).

This is the thinnest host over the HIP engine: one Python object per device
state vector, every call forwarded to the C entry point of the same name.
It fails loudly when the shared library is missing; there is no fallback.
"""
import ctypes
import os

import numpy as np

from .circuits import encode_observables
from .gates import GATE_ID

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpl_amd.so")

PL_C64, PL_C128 = 0, 1
_ERRORS = {1: ValueError, 2: IndexError, 3: RuntimeError}

# every symbol declared in include/pl_amd.h
EXPORTS = [
    "pl_last_error", "pl_device_info", "pl_sv_create", "pl_sv_destroy", "pl_sv_num_qubits",
    "pl_sv_dtype", "pl_sv_upload", "pl_sv_download", "pl_sv_copy", "pl_sv_reset",
    "pl_sv_set_basis_state", "pl_sv_device_ptr", "pl_sv_synchronize", "pl_sv_apply_gate",
    "pl_sv_apply_ops", "pl_sv_apply_matrix", "pl_sv_apply_generator", "pl_expval", "pl_var",
    "pl_expval_named", "pl_var_named", "pl_expval_matrix", "pl_probs", "pl_sample",
    "pl_generate_samples", "pl_adjoint_jacobian", "pl_timer_start", "pl_timer_stop",
    "pl_event_record", "pl_event_elapsed", "pl_sv_apply_observable", "pl_batch_create",
    "pl_batch_run", "pl_batch_info", "pl_batch_destroy", "pl_sv_apply_ops_fused", "pl_fuse_tape",
    "pl_sv_norm2", "pl_sv_distance", "pl_mlp_create", "pl_mlp_destroy", "pl_mlp_num_params",
    "pl_mlp_get_params", "pl_mlp_set_params", "pl_mlp_set_loss_weights", "pl_mlp_predict",
    "pl_mlp_gradients", "pl_mlp_load_data", "pl_mlp_train", "pl_vjp", "pl_expval_sparse",
    "pl_var_sparse", "pl_generate_mcmc_samples", "pl_sv_apply_ops_tiled", "pl_tile_plan",
    "pl_transfer_counters", "pl_expval_noalloc", "pl_sv_set_deferred", "pl_sv_deferred_count",
]

PL_FUSE_RENORM, PL_FUSE_NO_MFMA = 1, 2

_lib = None


def load():
    """Load libpl_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(LIB_PATH)
        lib.pl_last_error.restype = ctypes.c_char_p
        lib.pl_sv_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        for name in EXPORTS:
            fn = getattr(lib, name)
            if name != "pl_last_error":
                fn.restype = ctypes.c_int
        lib.pl_sv_set_basis_state.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        lib.pl_sample.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        lib.pl_generate_samples.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64,
                                            ctypes.c_void_p]
        lib.pl_sv_upload.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        lib.pl_sv_download.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        lib.pl_batch_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]
        lib.pl_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.pl_expval_sparse.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        lib.pl_var_sparse.argtypes = lib.pl_expval_sparse.argtypes
        lib.pl_generate_mcmc_samples.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                                 ctypes.c_int64, ctypes.c_uint64, ctypes.c_int,
                                                 ctypes.c_void_p]
        lib.pl_mlp_create.argtypes = [ctypes.c_int] * 5 + [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
        lib.pl_mlp_load_data.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        lib.pl_mlp_train.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_void_p]
        _lib = lib
    return _lib


def transfer_counters():
    """(bytes uploaded, bytes downloaded) by state uploads / downloads so far."""
    h2d, d2h = ctypes.c_uint64(), ctypes.c_uint64()
    _check(load().pl_transfer_counters(ctypes.byref(h2d), ctypes.byref(d2h)))
    return h2d.value, d2h.value


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        msg = load().pl_last_error().decode()
        raise _ERRORS.get(rc, RuntimeError)(msg)


def device_info():
    buf = ctypes.create_string_buffer(512)
    _check(load().pl_device_info(buf, 512))
    return buf.value.decode()


def _i32(x):
    a = np.ascontiguousarray(np.asarray(x, np.int32).reshape(-1))
    return a if a.size else np.zeros(1, np.int32)


def _f64(x):
    a = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))
    return a if a.size else np.zeros(1, np.float64)


def _cmat(m):
    return np.ascontiguousarray(np.asarray(m, np.complex128).reshape(-1)).view(np.float64)


class DeviceStateVector:
    """A state vector resident in HBM (pl_sv_t)."""

    def __init__(self, num_qubits, dtype=np.complex128):
        lib = load()
        self.dtype = np.dtype(dtype)
        code = PL_C128 if self.dtype == np.complex128 else PL_C64
        if self.dtype not in (np.complex128, np.complex64):
            raise TypeError("dtype must be complex64 or complex128")
        h = ctypes.c_void_p()
        _check(lib.pl_sv_create(int(num_qubits), code, ctypes.byref(h)))
        self.h = h
        self.num_qubits = int(num_qubits)

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr)
        n = int(arr.size).bit_length() - 1
        if arr.ndim != 1 or arr.size != 1 << n:
            raise ValueError("state must be a 1-D array of length 2^n")
        sv = cls(n, arr.dtype)
        sv.upload(arr)
        return sv

    def close(self):
        if getattr(self, "h", None):
            load().pl_sv_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def size(self):
        return 1 << self.num_qubits

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, self.dtype)
        _check(load().pl_sv_upload(self.h, _p(arr), arr.nbytes))

    def state(self):
        out = np.empty(self.size, self.dtype)
        _check(load().pl_sv_download(self.h, _p(out), out.nbytes))
        return out

    def device_ptr(self):
        """Device address of the amplitudes (owned by this object)."""
        out = ctypes.c_void_p()
        _check(load().pl_sv_device_ptr(self.h, ctypes.byref(out)))
        return out.value

    def copy_from(self, other):
        _check(load().pl_sv_copy(self.h, other.h))

    def reset(self):
        _check(load().pl_sv_reset(self.h))

    def set_basis_state(self, index):
        _check(load().pl_sv_set_basis_state(self.h, int(index)))

    def synchronize(self):
        _check(load().pl_sv_synchronize(self.h))

    def set_deferred(self, on=True):
        """Queue gates and apply them as tile passes when the state is next
        used (pl_sv_set_deferred); on=False applies the queue."""
        _check(load().pl_sv_set_deferred(self.h, int(bool(on))))

    def deferred_count(self):
        out = ctypes.c_int()
        _check(load().pl_sv_deferred_count(self.h, ctypes.byref(out)))
        return out.value

    # ---- gates ---------------------------------------------------------------
    def apply(self, name, wires, inverse=False, params=()):
        w, p = _i32(wires), _f64(params)
        _check(load().pl_sv_apply_gate(self.h, GATE_ID[name], _p(w), len(wires), int(inverse),
                                       _p(p), len(params)))

    def apply_tape(self, tape):
        ids, nw, wires, inv, npar, par = tape.encode()
        _check(load().pl_sv_apply_ops(self.h, len(ids), _p(_i32(ids)), _p(_i32(nw)),
                                      _p(_i32(wires)), _p(inv if inv.size else np.zeros(1, np.uint8)),
                                      _p(_i32(npar)), _p(_f64(par))))

    def apply_tape_fused(self, tape, max_wires=5, renorm=False, mfma=True):
        """Apply a tape as fused dense blocks (pl_sv_apply_ops_fused); returns
        the number of blocks."""
        ids, nw, wires, inv, npar, par = tape.encode()
        flags = (PL_FUSE_RENORM if renorm else 0) | (0 if mfma else PL_FUSE_NO_MFMA)
        nb = ctypes.c_int()
        _check(load().pl_sv_apply_ops_fused(
            self.h, len(ids), _p(_i32(ids)), _p(_i32(nw)), _p(_i32(wires)),
            _p(inv if inv.size else np.zeros(1, np.uint8)), _p(_i32(npar)), _p(_f64(par)),
            int(max_wires), flags, ctypes.byref(nb)))
        return nb.value

    def apply_tape_tiled(self, tape, tile_bits=0, low_bits=-1, renorm=False):
        """Apply a tape as tile passes (pl_sv_apply_ops_tiled); returns the
        number of launches.  tile_bits=0 / low_bits=-1: the defaults of the
        precision; renorm rescales to unit norm at the end (FP64 norm)."""
        ids, nw, wires, inv, npar, par = tape.encode()
        npass = ctypes.c_int()
        _check(load().pl_sv_apply_ops_tiled(
            self.h, len(ids), _p(_i32(ids)), _p(_i32(nw)), _p(_i32(wires)),
            _p(inv if inv.size else np.zeros(1, np.uint8)), _p(_i32(npar)), _p(_f64(par)),
            int(tile_bits), int(low_bits), PL_FUSE_RENORM if renorm else 0, ctypes.byref(npass)))
        return npass.value

    def norm2(self):
        out = ctypes.c_double()
        _check(load().pl_sv_norm2(self.h, ctypes.byref(out)))
        return out.value

    def distance(self, other):
        """||self - other||_2 in FP64 (the precisions may differ)."""
        out = ctypes.c_double()
        _check(load().pl_sv_distance(self.h, other.h, ctypes.byref(out)))
        return out.value

    def vjp(self, tape, dy, tp):
        """<dy| d psi / d theta_t> for the trainable parameters tp (pl_vjp)."""
        ids, nw, wires, inv, npar, par = tape.encode()
        dyv = np.ascontiguousarray(np.asarray(dy, np.complex128).reshape(-1))
        if dyv.size != self.size:
            raise ValueError("dy must have 2^n entries")
        out = np.zeros(max(1, len(tp)), np.complex128)
        _check(load().pl_vjp(self.h, len(ids), _p(_i32(ids)), _p(_i32(nw)), _p(_i32(wires)),
                             _p(inv if inv.size else np.zeros(1, np.uint8)), _p(_i32(npar)),
                             _p(_f64(par)), _p(dyv.view(np.float64)), len(tp), _p(_i32(tp)),
                             _p(out.view(np.float64))))
        return out[:len(tp)]

    def _sparse(self, fn, row_map, entries, values):
        rm = np.ascontiguousarray(np.asarray(row_map, np.int64))
        en = np.ascontiguousarray(np.asarray(entries, np.int64))
        va = np.ascontiguousarray(np.asarray(values, np.complex128))
        if en.size != va.size:
            raise ValueError("entries and values differ in length")
        out = ctypes.c_double()
        _check(fn(self.h, _p(rm), rm.size, _p(en if en.size else np.zeros(1, np.int64)),
                  _p((va if va.size else np.zeros(1, np.complex128)).view(np.float64)), va.size,
                  ctypes.byref(out)))
        return out.value

    def expval_sparse(self, row_map, entries, values):
        return self._sparse(load().pl_expval_sparse, row_map, entries, values)

    def var_sparse(self, row_map, entries, values):
        return self._sparse(load().pl_var_sparse, row_map, entries, values)

    def generate_mcmc_samples(self, num_shots, kernel=0, num_burnin=100, seed=0, nchains=0):
        """Metropolis-Hastings samples (pl_generate_mcmc_samples): (shots, n) bits."""
        out = np.zeros((int(num_shots), self.num_qubits), np.uint64)
        _check(load().pl_generate_mcmc_samples(self.h, int(kernel), int(num_burnin), int(num_shots),
                                               int(seed), int(nchains), _p(out)))
        return out

    def apply_matrix(self, matrix, wires, inverse=False):
        m, w = _cmat(matrix), _i32(wires)
        _check(load().pl_sv_apply_matrix(self.h, _p(m), _p(w), len(wires), int(inverse)))

    def apply_generator(self, name, wires, adj=False):
        w = _i32(wires)
        s = ctypes.c_double()
        _check(load().pl_sv_apply_generator(self.h, GATE_ID[name], _p(w), len(wires), int(adj),
                                            ctypes.byref(s)))
        return s.value

    # ---- measurements ----------------------------------------------------------
    def _obs(self, ob):
        prog, coeffs, mats = encode_observables([ob])
        return prog, coeffs, np.ascontiguousarray(mats).view(np.float64)

    def apply_observable(self, ob):
        prog, coeffs, mats = self._obs(ob)
        _check(load().pl_sv_apply_observable(self.h, _p(prog), _p(coeffs), _p(mats)))

    def expval(self, ob):
        prog, coeffs, mats = self._obs(ob)
        out = ctypes.c_double()
        _check(load().pl_expval(self.h, _p(prog), _p(coeffs), _p(mats), ctypes.byref(out)))
        return out.value

    def expval_noalloc(self, ob):
        """pl_expval_noalloc: Pauli-word observables, no device allocation."""
        prog, coeffs, mats = self._obs(ob)
        out = ctypes.c_double()
        _check(load().pl_expval_noalloc(self.h, _p(prog), _p(coeffs), _p(mats), ctypes.byref(out)))
        return out.value

    def var(self, ob):
        prog, coeffs, mats = self._obs(ob)
        out = ctypes.c_double()
        _check(load().pl_var(self.h, _p(prog), _p(coeffs), _p(mats), ctypes.byref(out)))
        return out.value

    def expval_named(self, name, wires, var=False):
        w = _i32(wires)
        out = ctypes.c_double()
        fn = load().pl_var_named if var else load().pl_expval_named
        _check(fn(self.h, GATE_ID[name], _p(w), len(wires), ctypes.byref(out)))
        return out.value

    def expval_matrix(self, matrix, wires):
        m, w = _cmat(matrix), _i32(wires)
        out = ctypes.c_double()
        _check(load().pl_expval_matrix(self.h, _p(m), _p(w), len(wires), ctypes.byref(out)))
        return out.value

    def probs(self, wires=None):
        if wires is None:
            out = np.empty(self.size, np.float64)
            _check(load().pl_probs(self.h, _p(np.zeros(1, np.int32)), -1, _p(out)))
        else:
            out = np.empty(1 << len(wires), np.float64)
            _check(load().pl_probs(self.h, _p(_i32(wires)), len(wires), _p(out)))
        return out

    def sample(self, uniforms):
        u = _f64(uniforms)
        idx = np.empty(max(1, len(uniforms)), np.int64)
        _check(load().pl_sample(self.h, _p(u), len(uniforms), _p(idx)))
        return idx[: len(uniforms)]

    def generate_samples(self, num_samples, seed):
        out = np.empty((max(1, num_samples), self.num_qubits), np.uint64)
        _check(load().pl_generate_samples(self.h, int(seed) & ((1 << 64) - 1), int(num_samples),
                                          _p(out)))
        return out[:num_samples]

    def adjoint_jacobian(self, tape, obs_list, tp):
        ids, nw, wires, inv, npar, par = tape.encode()
        prog, coeffs, mats = encode_observables(obs_list)
        mats = np.ascontiguousarray(mats).view(np.float64)
        jac = np.zeros(max(1, len(obs_list) * len(tp)), np.float64)
        _check(load().pl_adjoint_jacobian(
            self.h, len(ids), _p(_i32(ids)), _p(_i32(nw)), _p(_i32(wires)),
            _p(inv if inv.size else np.zeros(1, np.uint8)), _p(_i32(npar)), _p(_f64(par)),
            _p(prog), _p(coeffs), _p(mats), len(tp), _p(_i32(tp)), _p(jac)))
        return jac[: len(obs_list) * len(tp)].reshape(len(obs_list), len(tp))

    # ---- timing ----------------------------------------------------------------
    def timer_start(self):
        _check(load().pl_timer_start(self.h))

    def timer_stop(self):
        ms = ctypes.c_double()
        _check(load().pl_timer_stop(self.h, ctypes.byref(ms)))
        return ms.value

    def event_record(self, slot):
        _check(load().pl_event_record(self.h, int(slot)))

    def event_elapsed(self, a, b):
        ms = ctypes.c_double()
        _check(load().pl_event_elapsed(self.h, int(a), int(b), ctypes.byref(ms)))
        return ms.value


def fuse_tape(num_qubits, tape, max_wires=5, pad_to=0):
    """Host-only fusion plan (pl_fuse_tape): list of blocks
    (tape op indices in application order, bits or None for an unfused op,
    matrix or None)."""
    ids, nw, wires, inv, npar, par = tape.encode()
    cap = max(1, len(ids))
    nb = ctypes.c_int()
    first, nops, nbits = (np.zeros(cap, np.int32) for _ in range(3))
    order = np.zeros(cap, np.int32)
    bits = np.zeros(8 * cap, np.int32)
    mats = np.zeros(2048 * cap, np.float64)
    _check(load().pl_fuse_tape(int(num_qubits), len(ids), _p(_i32(ids)), _p(_i32(nw)), _p(_i32(wires)),
                               _p(inv if inv.size else np.zeros(1, np.uint8)), _p(_i32(npar)),
                               _p(_f64(par)), int(max_wires), int(pad_to), cap, ctypes.byref(nb),
                               _p(first), _p(nops), _p(order), _p(nbits), _p(bits), _p(mats)))
    out = []
    for b in range(nb.value):
        k = int(nbits[b])
        ops = order[first[b]:first[b] + nops[b]].tolist()
        if k == 0:
            out.append((ops, None, None))
            continue
        d = 1 << k
        m = mats[2048 * b: 2048 * b + 2 * d * d].view(np.complex128).reshape(d, d).copy()
        out.append((ops, bits[8 * b: 8 * b + k].tolist(), m))
    return out


def tile_plan(num_qubits, tape, tile_bits=12, low_bits=6):
    """Host-only pass plan of pl_sv_apply_ops_tiled: list of passes
    (tiled, high tile bits, tape op indices in application order)."""
    ids, nw, wires, inv, npar, par = tape.encode()
    cap = max(1, len(ids))
    npass = ctypes.c_int()
    pn, pt = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    pb = np.zeros(16 * cap, np.int32)
    order = np.zeros(cap, np.int32)
    _check(load().pl_tile_plan(int(num_qubits), len(ids), _p(_i32(ids)), _p(_i32(nw)), _p(_i32(wires)),
                               _p(inv if inv.size else np.zeros(1, np.uint8)), _p(_i32(npar)),
                               _p(_f64(par)), int(tile_bits), int(low_bits), cap, ctypes.byref(npass),
                               _p(pn), _p(pt), _p(pb), _p(order)))
    out, at = [], 0
    for p in range(npass.value):
        hb = [int(b) for b in pb[16 * p: 16 * p + 16] if b >= 0]
        out.append((bool(pt[p]), hb, order[at: at + pn[p]].tolist()))
        at += int(pn[p])
    return out


class BatchPlan:
    """Batched layered-ansatz circuits (pl_batch_*): one hipGraph replay gives
    every circuit's energy and gradient."""

    def __init__(self, nqubits, depths, pattern, entangler, observables, dtype=np.complex128,
                 tile_width=0):
        lib = load()
        self.nqubits = [int(x) for x in nqubits]
        self.depths = [int(x) for x in depths]
        if len(self.nqubits) != len(self.depths) or len(observables) != len(self.nqubits):
            raise ValueError("need one qubit count, depth and observable per circuit")
        from .circuits import ENTANGLERS
        ent = ENTANGLERS[entangler] if isinstance(entangler, str) else int(entangler)
        prog, coeffs, _ = encode_observables(observables)
        code = PL_C128 if np.dtype(dtype) == np.complex128 else PL_C64
        h = ctypes.c_void_p()
        _check(lib.pl_batch_create(code, len(self.nqubits), _p(_i32(self.nqubits)),
                                   _p(_i32(self.depths)), pattern.encode(), ent, _p(prog),
                                   _p(coeffs), int(tile_width), ctypes.byref(h)))
        self.h = h
        w, nl, npar = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        _check(lib.pl_batch_info(self.h, ctypes.byref(w), ctypes.byref(nl), ctypes.byref(npar)))
        self.tile_width, self.launches, self.nparams = w.value, nl.value, npar.value
        self.pattern = pattern
        self.param_offsets = np.concatenate(
            [[0], np.cumsum([d * n * len(pattern) for n, d in zip(self.nqubits, self.depths)])])

    def run(self, params, grads=True):
        p = _f64(params)
        if p.size < self.nparams:
            raise ValueError("need one angle per rotation")
        e = np.zeros(len(self.nqubits))
        g = np.zeros(max(1, self.nparams)) if grads else None
        _check(load().pl_batch_run(self.h, _p(p), _p(e), _p(g) if grads else None))
        return (e, g[: self.nparams]) if grads else e

    def split(self, flat):
        o = self.param_offsets
        return [flat[o[j]:o[j + 1]] for j in range(len(self.nqubits))]

    def close(self):
        if getattr(self, "h", None):
            load().pl_batch_destroy(self.h)
            self.h = None

    __del__ = close


def _f32(x):
    return np.ascontiguousarray(np.asarray(x, np.float32))


class Mlp:
    """The surrogate network (pl_mlp_*): d0 -> d1 -> d2 -> d3, GELU hidden
    layers, BF16 matrix cores with FP32 accumulation, FP32 Adam."""

    def __init__(self, dims=(192, 512, 512, 193), max_batch=1024, seed=1005):
        self.dims = tuple(int(d) for d in dims)
        self.max_batch = int(max_batch)
        h = ctypes.c_void_p()
        _check(load().pl_mlp_create(*self.dims, self.max_batch, int(seed), ctypes.byref(h)))
        self.h = h
        n = ctypes.c_int64()
        _check(load().pl_mlp_num_params(self.h, ctypes.byref(n)))
        self.nparams = n.value

    def close(self):
        if getattr(self, "h", None):
            load().pl_mlp_destroy(self.h)
            self.h = None

    __del__ = close

    def params(self):
        out = np.empty(self.nparams, np.float32)
        _check(load().pl_mlp_get_params(self.h, _p(out)))
        return out

    def set_params(self, p):
        p = _f32(p)
        if p.size != self.nparams:
            raise ValueError("wrong parameter count")
        _check(load().pl_mlp_set_params(self.h, _p(p)))

    def unpack(self, flat):
        """-> [(W1, b1), (W2, b2), (W3, b3)] views of a flat parameter vector."""
        d, out, o = self.dims, [], 0
        for l in range(3):
            w = flat[o:o + d[l] * d[l + 1]].reshape(d[l], d[l + 1])
            o += d[l] * d[l + 1]
            b = flat[o:o + d[l + 1]]
            o += d[l + 1]
            out.append((w, b))
        return out

    def set_loss_weights(self, w):
        w = _f32(w)
        if w.size != self.dims[3]:
            raise ValueError("one weight per output")
        _check(load().pl_mlp_set_loss_weights(self.h, _p(w)))

    def predict(self, x):
        x = _f32(x).reshape(-1, self.dims[0])
        y = np.empty((x.shape[0], self.dims[3]), np.float32)
        for r in range(0, x.shape[0], self.max_batch):
            xr = np.ascontiguousarray(x[r:r + self.max_batch])
            yr = np.empty((xr.shape[0], self.dims[3]), np.float32)
            _check(load().pl_mlp_predict(self.h, _p(xr), xr.shape[0], _p(yr)))
            y[r:r + xr.shape[0]] = yr
        return y

    def gradients(self, x, t):
        x, t = _f32(x), _f32(t)
        g = np.empty(self.nparams, np.float32)
        loss = ctypes.c_double()
        _check(load().pl_mlp_gradients(self.h, _p(x), _p(t), x.shape[0], ctypes.byref(loss), _p(g)))
        return loss.value, g

    def load_data(self, x, t):
        x, t = _f32(x), _f32(t)
        _check(load().pl_mlp_load_data(self.h, _p(x), _p(t), x.shape[0]))

    def train(self, steps, batch, lr):
        loss = ctypes.c_double()
        _check(load().pl_mlp_train(self.h, int(steps), int(batch), float(lr), ctypes.byref(loss)))
        return loss.value
