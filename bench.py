"""Benchmark: gate-application HBM throughput on one MI355X.

Workload (one "step"): the seeded 50-layer random circuit on 30 qubits in
complex128 (16 GiB state, 990 gates drawn from {Hadamard, T, RX, RY, RZ, Rot,
CRX, CRZ, IsingXX, IsingZZ, SWAP, Toffoli}), applied gate by gate to a state
resident in HBM.  value = algorithmic bytes moved by the gates / device time,
where a gate that can change a fraction f of the amplitudes moves
2 * f * 2^n * 16 bytes.  Every launch is bracketed by HIP events on the
engine's stream (the stream the kernels run on).  Also reported, each in its
own object of the JSON line: the 100-term Hamiltonian expectation and 10^6
inverse-CDF samples on the final state (event-timed, GB/s of their
algorithmic bytes), the same circuit through the tile-pass executor, the
20-qubit adjoint Jacobian, the VJP, sparse-Hamiltonian and Metropolis entry
points (next_rows), the batching form (4096 circuits, forward + adjoint), the
precision form (33 qubits, complex64), the learned surrogate, and a CPU
baseline (the reference's Lightning-Qubit kernels compiled from source, or the
oracle restatement) on a bounded sample of the same gates.

Modes for single measurements: copy (the copy kernel that calibrates the
traffic counters), stream (the device's streaming ceilings), measure (<H>
and sampling on the circuit's state), batch, precision, surrogate, c1, next.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--qubits N] [--mode M]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from lightning_unification_amd.circuits import (random_hamiltonian, random_layers,  # noqa: E402
                                                sample_uniforms, Tape)
from lightning_unification_amd.gates import touched_fraction  # noqa: E402

HBM_PEAK_GBS = 8000.0

DIAGONAL = {"RZ", "CRZ", "IsingZZ"}  # lowered with a diagonal matrix (the DIAG instantiations)


def kernel_for(name, wires, n):
    """Kernel instantiation a gate of this workload runs as (the dispatch of
    Engine::launch_small in csrc/pl_amd.hip)."""
    bits = [n - 1 - w for w in wires]
    m = 1 if name == "T" else (4 if name in ("IsingXX", "IsingZZ") else 2)
    diag = "true" if name in DIAGONAL else "false"
    if m == 2 and name != "SWAP":
        target, others = bits[-1], bits[:-1]  # 1q gates, CRX/CRZ, Toffoli
        if target < 6 and all(b >= 6 for b in others) and n - len(others) >= 6:
            return f"k_pair_lane<double,{diag}>"
    nhi = sum(1 for b in bits if b >= 8)
    if min(bits) < 3 and nhi <= 3 and n >= 8 + nhi:
        return f"k_group_tile<double,{m},{nhi},{diag}>"
    return f"k_group_flat<double,{m},{diag}>"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` as measured by rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs (scaled by a copy-kernel calibration), read from the
    committed summary under profiles/; None when no summary covers it."""
    pdir = os.path.join(ROOT, "profiles")
    key = f"traffic_bytes_per_launch {kernel}:"
    for name in sorted(os.listdir(pdir)) if os.path.isdir(pdir) else []:
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                if line.startswith(key):
                    return int(line.split(":")[1]), f"profiles/{name}"
    return None, None


def cpu_baseline(tape, n, budget_s=20.0):
    """Time the reference LQ kernels (or the oracle) on the first gates of the
    same circuit at the same size; single-threaded like the reference's gate
    kernels (no OpenMP under simulators/lightning_qubit/gates)."""
    try:
        from oracle.pyoracle import Oracle, REF_SO
        kind = "reference" if os.path.exists(REF_SO) else "port"
        orc = Oracle(ref=(kind == "reference"))
    except Exception as e:  # oracle not built
        log(f"cpu baseline unavailable: {e}")
        return None
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    sv = np.zeros(1 << n, np.complex128)
    sv[0] = 1
    done, nbytes, t_total = 0, 0.0, 0.0
    for op in tape.ops:
        t = Tape()
        t.ops.append(op)
        t0 = time.perf_counter()
        orc.apply_tape(sv, t)
        dt = time.perf_counter() - t0
        t_total += dt
        nbytes += 2 * touched_fraction(op[0]) * (1 << n) * 16
        done += 1
        if t_total > budget_s:
            break
    del sv
    return {"value": round(nbytes / t_total / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": kind,
            "sample": f"first {done} gates of the same {n}-qubit complex128 circuit, "
                      f"{t_total:.1f} s single-threaded",
            **host_info()}


def host_info():
    """CPU model, memory and hardware threads of the host the CPU baseline ran on."""
    model, mem_kib = None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal"):
                    mem_kib = int(line.split()[1])
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "ram_GiB": round(mem_kib / 2 ** 20, 1) if mem_kib else None,
            "host_threads": os.cpu_count(), "threads_usable": usable}


def bench_batch(count, plain_budget_s=8.0, reps=3, cpu_budget_s=10.0):
    """Batching form: `count` HEA circuits (8..20 qubits, depth 4..12, RY+RZ
    + CNOT ring, periodic TFIM each), energy + full gradient for every circuit.
    Batched: one hipGraph replay of pl_batch_run per step (all circuits).
    Plain: the same circuits one at a time through the single-state path
    (reset, apply_ops, expval, pl_adjoint_jacobian), on a bounded sample.
    grads/s := complete (E, grad E) circuits per second."""
    from lightning_unification_amd import capi
    from lightning_unification_amd.circuits import batch_shapes, layered_tape, tfim

    ns, ds, hs, params = batch_shapes(count)
    obs = [tfim(n, 1.0, h) for n, h in zip(ns, hs)]
    out = {"circuits": count, "params": int(params.size),
           "qubits": [int(min(ns)), int(max(ns))], "depths": [int(min(ds)), int(max(ds))]}
    res = {}
    for name, dt in (("c128", np.complex128), ("c64", np.complex64)):
        t0 = time.perf_counter()
        plan = capi.BatchPlan(ns, ds, "YZ", "ring", obs, dtype=dt)
        t_setup = time.perf_counter() - t0
        e, g = plan.run(params)  # warm replay
        t0 = time.perf_counter()
        for _ in range(reps):
            e, g = plan.run(params)
        dt_s = (time.perf_counter() - t0) / reps
        res[name] = (e, g)
        out[name] = {"grads_per_s": round(count / dt_s, 1), "ms_per_batch": round(dt_s * 1e3, 2),
                     "tile_width": plan.tile_width, "graph_launches": plan.launches,
                     "setup_s": round(t_setup, 2)}
        log(f"batched {name}: {dt_s*1e3:.1f} ms per {count} circuits (W={plan.tile_width})")
        plan.close()
    e128, g128 = res["c128"]
    e64, g64 = res["c64"]
    out["c64_vs_c128"] = {"max_abs_energy_err": float(np.max(np.abs(e64 - e128))),
                          "max_abs_grad_err": float(np.max(np.abs(g64 - g128)))}

    # plain: the single-circuit API, circuit by circuit over a seeded random
    # sample of the same batch (an unbiased mix), until the budget is spent
    from lightning_unification_amd.circuits import Rng
    offs = np.concatenate([[0], np.cumsum([2 * n * d for n, d in zip(ns, ds)])])
    done, t_plain, max_diff = 0, 0.0, 0.0
    svs = {}
    for j in Rng(5).shuffle(range(count)):
        n, d = ns[j], ds[j]
        p = params[offs[j]:offs[j + 1]]
        tape = layered_tape(n, d, "YZ", "ring", p)
        sv = svs.get(n) or svs.setdefault(n, capi.DeviceStateVector(n, np.complex128))
        t0 = time.perf_counter()
        sv.reset()
        sv.apply_tape(tape)
        ej = sv.expval(obs[j])
        jac = sv.adjoint_jacobian(tape, [obs[j]], list(range(p.size)))
        t_plain += time.perf_counter() - t0
        max_diff = max(max_diff, abs(ej - e128[j]),
                       float(np.max(np.abs(jac[0] - g128[offs[j]:offs[j + 1]]))))
        done += 1
        if t_plain > plain_budget_s:
            break
    for sv in svs.values():
        sv.close()
    plain_gps = done / t_plain
    out["plain_c128"] = {"grads_per_s": round(plain_gps, 1), "sample_circuits": done,
                         "sample_s": round(t_plain, 2), "max_abs_diff_vs_batched": max_diff}
    out["speedup_c128"] = round(out["c128"]["grads_per_s"] / plain_gps, 1)
    out["speedup_c64"] = round(out["c64"]["grads_per_s"] / plain_gps, 1)
    log(f"plain: {done} circuits in {t_plain:.1f} s; speedup c128 {out['speedup_c128']}x")

    # the reference's CPU path on the same seeded sample (forward, expval,
    # adjoint with one observable), bounded by cpu_budget_s
    try:
        from oracle.pyoracle import Oracle, REF_SO
        kind = "reference" if os.path.exists(REF_SO) else "port"
        orc = Oracle(ref=(kind == "reference"))
        done_c, t_cpu = 0, 0.0
        for j in Rng(5).shuffle(range(count)):
            n, d = ns[j], ds[j]
            p = params[offs[j]:offs[j + 1]]
            tape = layered_tape(n, d, "YZ", "ring", p)
            t0 = time.perf_counter()
            s = np.zeros(1 << n, np.complex128)
            s[0] = 1
            orc.apply_tape(s, tape)
            orc.expval(s, obs[j])
            orc.adjoint(s, tape, [obs[j]], list(range(p.size)))
            t_cpu += time.perf_counter() - t0
            done_c += 1
            if t_cpu > cpu_budget_s:
                break
        out["cpu_reference_c128"] = {"grads_per_s": round(done_c / t_cpu, 2), "sample_circuits": done_c,
                                     "sample_s": round(t_cpu, 2), "kind": kind,
                                     "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))}
        log(f"cpu reference: {done_c} circuits in {t_cpu:.1f} s")
    except Exception as e:  # oracle not built
        log(f"batch cpu leg unavailable: {e}")
    return out


def bench_c1():
    """Plain method at the reference's CPU-runnable size: 20 qubits, 10 layers
    of RX, RY, RZ per wire + CNOT ladder, <Z_w> for every wire and the full
    20 x 600 adjoint Jacobian (complex128).  Reports the GPU time of forward +
    expvals + Jacobian, the unfused-minimum algorithmic bytes of the Jacobian
    (sum over ops of (K+1) 2 f S + per trainable op (K+1) S, plus K S) and the
    bytes this engine's fused adjoint moves ((K+G) 2 f S per op for G groups
    of observables, K + G state preparations), and the compiled reference (all host cores, OpenMP over
    observables as in the reference) beside the GPU on a one-layer prefix."""
    from lightning_unification_amd import capi
    from lightning_unification_amd.circuits import plain_layers
    n, K = 20, 20
    S = (1 << n) * 16
    tape, obs, tp = plain_layers(n, 10, seed=1001)
    sv = capi.DeviceStateVector(n, np.complex128)

    def run(t, params):
        sv.reset()
        sv.apply_tape(t)
        e = [sv.expval(o) for o in obs]
        return e, sv.adjoint_jacobian(t, obs, params)

    run(tape, tp)  # warm
    t0 = time.perf_counter()
    _, jac = run(tape, tp)
    t_gpu = time.perf_counter() - t0
    alg = sum((K + 1) * 2 * touched_fraction(op[0]) * S for op in tape.ops) + len(tp) * (K + 1) * S + K * S
    # the engine runs the sweep over groups of observables that fit its 192 MiB
    # cache budget, each group with its own lambda (two groups of 10 here)
    budget = 192 << 20
    groups = 1
    if S * (K + 1) > budget and 2 * S <= budget:
        fit = max(1, budget // S - 1)
        groups = -(-K // fit)
    moved = sum((K + groups) * 2 * touched_fraction(op[0]) * S for op in tape.ops) + (K + groups) * 2 * S
    out = {"qubits": n, "ops": len(tape), "params": len(tp), "observables": K,
           "gpu_forward_expval_jacobian_ms": round(t_gpu * 1e3, 1),
           "jacobian_algorithmic_GB": round(alg / 1e9, 1),
           "jacobian_moved_GB": round(moved / 1e9, 1),
           "algorithmic_GBs": round(alg / t_gpu / 1e9, 1)}
    # one-layer prefix on both sides
    pre, _, _ = plain_layers(n, 1, seed=1001)
    ptp = list(range(pre.num_params()))
    t0 = time.perf_counter()
    run(pre, ptp)
    t_gpu1 = time.perf_counter() - t0
    try:
        from oracle.pyoracle import Oracle, REF_SO
        kind = "reference" if os.path.exists(REF_SO) else "port"
        orc = Oracle(ref=(kind == "reference"))
        s = np.zeros(1 << n, np.complex128)
        s[0] = 1
        t0 = time.perf_counter()
        orc.apply_tape(s, pre)
        ref_e = [orc.expval(s, o) for o in obs]
        ref_j = orc.adjoint(s, pre, obs, ptp)
        t_cpu1 = time.perf_counter() - t0
        _, gj = run(pre, ptp)
        out["one_layer_prefix"] = {"ops": len(pre), "params": len(ptp), "gpu_ms": round(t_gpu1 * 1e3, 1),
                                   "cpu_ms": round(t_cpu1 * 1e3, 1), "cpu_kind": kind,
                                   "cpu_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
                                   "max_abs_jacobian_diff": float(np.max(np.abs(gj - ref_j)))}
        del ref_e
    except Exception as e:  # oracle not built
        log(f"c1 cpu leg unavailable: {e}")
    sv.close()
    log(f"C1: forward + 20 expvals + 20x600 Jacobian in {t_gpu*1e3:.1f} ms")
    return out


def bench_next():
    """The entry points beside the path, each timed end to end through the C ABI
    on the state of the 20-qubit, 10-layer plain circuit (complex128):
    - VJP (pl_vjp) of the whole tape against a seeded cotangent dy.
      Algorithmic bytes: two states (psi, mu) swept backwards, 2 x 2 f S per
      op, plus one read of each state per trainable parameter for the inner
      product.
    - <H> and Var(H) of a seeded sparse Hermitian H (pl_expval_sparse /
      pl_var_sparse) at 22 qubits, 5 non-zeros per row (the diagonal and four
      bit-flip partners).  The boundary takes host CSR arrays, as the
      reference's does, so the time includes the PCIe upload of the 24 B per
      non-zero; the kernel alone reads those bytes plus psi once and checks
      the indices as it reads them.
    - Whole-state transfers (pl_sv_upload / pl_sv_download) of a 28-qubit
      complex128 state (4 GiB) from and to pageable host memory: the
      PCIe-inclusive rate of the boundary's copies, never the headline.
    - Metropolis samples (pl_generate_mcmc_samples), 10^5 shots after 100
      burn-in steps, with both transition kernels."""
    from lightning_unification_amd import capi
    from lightning_unification_amd.circuits import plain_layers
    out = {}
    n = 20
    S = (1 << n) * 16
    tape, _, tp = plain_layers(n, 10, seed=1001)
    sv = capi.DeviceStateVector(n, np.complex128)
    sv.apply_tape(tape)
    rng = np.random.default_rng(1003)
    dy = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    dy /= np.linalg.norm(dy)
    sv.vjp(tape, dy, tp)  # warm
    t0 = time.perf_counter()
    sv.vjp(tape, dy, tp)
    t_vjp = time.perf_counter() - t0
    alg = sum(2 * 2 * touched_fraction(op[0]) * S for op in tape.ops) + len(tp) * 2 * S
    out["vjp"] = {"qubits": n, "ops": len(tape), "params": len(tp), "ms": round(t_vjp * 1e3, 2),
                  "algorithmic_GB": round(alg / 1e9, 2), "algorithmic_GBs": round(alg / t_vjp / 1e9, 1)}
    log(f"VJP: {len(tape)} ops, {len(tp)} params in {t_vjp*1e3:.1f} ms")

    for kernel in (0, 1):
        sv.generate_mcmc_samples(1000, kernel=kernel, num_burnin=100, seed=7)  # warm
        t0 = time.perf_counter()
        bits = sv.generate_mcmc_samples(100_000, kernel=kernel, num_burnin=100, seed=7)
        t_mc = time.perf_counter() - t0
        out[f"mcmc_kernel{kernel}"] = {"qubits": n, "shots": int(bits.shape[0]), "burnin": 100,
                                       "ms": round(t_mc * 1e3, 2)}
        log(f"MCMC kernel {kernel}: 1e5 shots in {t_mc*1e3:.1f} ms")
    sv.close()

    ns = 22
    dim = 1 << ns
    masks = np.array([0, 1, 1 << 7, 1 << 13, 1 << 21], np.int64)
    coef = np.array([0.0, 0.5, -0.25, 0.75, 0.125])
    rows = np.arange(dim, dtype=np.int64)
    cols = np.sort(rows[:, None] ^ masks[None, :], axis=1)
    diag = rng.standard_normal(dim)
    flip = cols ^ rows[:, None]
    vals = np.where(flip == 0, diag[:, None], 0.0)
    for m, c in zip(masks[1:], coef[1:]):
        vals = np.where(flip == m, c, vals)
    entries = cols.reshape(-1)
    values = vals.reshape(-1).astype(np.complex128)
    row_map = np.arange(dim + 1, dtype=np.int64) * len(masks)
    svs = capi.DeviceStateVector(ns, np.complex128)
    svs.apply_tape(random_layers(ns, 4, seed=1004))
    svs.expval_sparse(row_map, entries, values)  # warm
    t0 = time.perf_counter()
    e = svs.expval_sparse(row_map, entries, values)
    t_e = time.perf_counter() - t0
    t0 = time.perf_counter()
    v = svs.var_sparse(row_map, entries, values)
    t_v = time.perf_counter() - t0
    svs.close()
    csr_bytes = entries.size * 24 + row_map.size * 8
    out["sparse_expval_var"] = {"qubits": ns, "nnz": int(entries.size), "expval": e, "var": v,
                                "expval_ms": round(t_e * 1e3, 2), "var_ms": round(t_v * 1e3, 2),
                                "csr_GB": round(csr_bytes / 1e9, 3),
                                "includes": "PCIe upload of the CSR arrays"}
    log(f"sparse H at {ns} qubits, nnz {entries.size}: <H> {t_e*1e3:.1f} ms, var {t_v*1e3:.1f} ms")
    svs.close()
    del row_map, entries, values

    nt = 28
    host = np.zeros(1 << nt, np.complex128)
    host.real[::4097] = 1.0
    with capi.DeviceStateVector(nt, np.complex128) as svt:
        svt.upload(host)  # warm: first use of the pinned slots
        t0 = time.perf_counter()
        for _ in range(2):
            svt.upload(host)
        t_up = (time.perf_counter() - t0) / 2
        back = svt.state()
        t0 = time.perf_counter()
        for _ in range(2):
            back = svt.state()
        t_down = (time.perf_counter() - t0) / 2
        assert np.array_equal(back, host)
    out["state_transfer"] = {"qubits": nt, "GB": round(host.nbytes / 1e9, 3),
                             "upload_ms": round(t_up * 1e3, 1), "download_ms": round(t_down * 1e3, 1),
                             "upload_GBs": round(host.nbytes / t_up / 1e9, 1),
                             "download_GBs": round(host.nbytes / t_down / 1e9, 1),
                             "host": "pageable"}
    log(f"state transfer, {nt} qubits: upload {t_up*1e3:.0f} ms, download {t_down*1e3:.0f} ms")
    return out


def bench_tiled(sv_ref, tape, n, reps):
    """The same circuit through the tile-pass executor (pl_sv_apply_ops_tiled,
    default tile shape): `reps` applications to a fresh |0>, each timed with
    HIP events on the engine's stream; then ||psi_tiled - psi_per_gate||_2
    against sv_ref, which holds the per-gate path after as many applications.
    hbm_GBs counts one read and one write of the state per pass (the bytes a
    pass must move); gate_equivalent_GBs the per-gate algorithmic bytes of the
    circuit over the same time."""
    from lightning_unification_amd import capi
    S = (1 << n) * 16
    gate_bytes = sum(2 * touched_fraction(op[0]) * S for op in tape.ops)
    sv = capi.DeviceStateVector(n, np.complex128)
    times, npass = [], 0
    for r in range(reps):
        sv.synchronize()
        sv.event_record(0)
        npass = sv.apply_tape_tiled(tape)
        sv.event_record(1)
        sv.synchronize()
        times.append(sv.event_elapsed(0, 1))
    dist = sv.distance(sv_ref)
    # the same gates one pl_sv_apply_gate call each, deferred (what a caller
    # applying gate by gate, like the drop-in module, gets): from the first
    # call to the queue applied, once, on a fresh |0>
    sv.reset()
    sv.set_deferred(True)
    sv.synchronize()
    sv.event_record(0)
    for name, wires, inv, params in tape.ops:
        sv.apply(name, wires, inv, params)
    sv.event_record(1)  # applies the queue first
    ms_deferred = sv.event_elapsed(0, 1)
    sv.set_deferred(False)
    sv.close()
    ms = min(times)
    out = {"passes": npass, "gates": len(tape), "ms_per_circuit": round(ms, 2),
           "ms_mean": round(sum(times) / len(times), 2),
           "hbm_GBs": round(npass * 2 * S / (ms / 1e3) / 1e9, 1),
           "hbm_frac": round(npass * 2 * S / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
           "gate_equivalent_GBs": round(gate_bytes / (ms / 1e3) / 1e9, 1),
           "l2_vs_per_gate": dist,
           "deferred_gate_calls": {"calls": len(tape.ops), "ms": round(ms_deferred, 2)}}
    log(f"tiled: {npass} passes, {ms:.1f} ms per circuit, |dpsi| {dist:.2e}; "
        f"{len(tape.ops)} deferred gate calls {ms_deferred:.1f} ms")
    return out


def distinct_xmasks(ham, n):
    """Number of distinct X/Y masks among the Hamiltonian's Pauli words (the
    Z-only words share the mask 0): one read of the state each is the
    algorithmic traffic of the expectation."""
    masks = set()
    for t in ham.terms:
        fs = t.factors if t.kind == "tensor" else [t]
        m = 0
        for f in fs:
            if f.name in ("PauliX", "PauliY"):
                m |= 1 << (n - 1 - f.wires[0])
        masks.add(m)
    return len(masks)


def bench_measurements(sv, ham, uniforms, n, reps=3):
    """<H> of the 100-term Hamiltonian, inverse-CDF sampling of the
    uniforms and the marginal probabilities of 8 wires on the state in sv:
    one warm call each, then `reps` calls timed
    with HIP events on the engine's stream (the best is reported) and by the
    wall clock (host checks, table staging and the copy of the indices
    back included).  Algorithmic bytes: one read of the state per distinct
    x-mask for <H>; one read of the state plus 16 bytes per sample for the
    sampler; one read of the state for the probabilities."""
    S = (1 << n) * 16
    out = {}
    e = sv.expval(ham)
    dev, wall = [], []
    for _ in range(reps):
        sv.synchronize()
        t0 = time.perf_counter()
        sv.event_record(0)
        e = sv.expval(ham)
        sv.event_record(1)
        wall.append((time.perf_counter() - t0) * 1e3)
        dev.append(sv.event_elapsed(0, 1))
    nm = distinct_xmasks(ham, n)
    ms = min(dev)
    out["expval_100_terms"] = {"value": e, "ms": round(ms, 2), "wall_ms": round(min(wall), 2),
                               "distinct_xmasks": nm, "algorithmic_GB": round(nm * S / 1e9, 2),
                               "GBs": round(nm * S / (ms / 1e3) / 1e9, 1),
                               "hbm_frac": round(nm * S / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    idx = sv.sample(uniforms)
    dev, wall = [], []
    for _ in range(reps):
        sv.synchronize()
        t0 = time.perf_counter()
        sv.event_record(0)
        idx = sv.sample(uniforms)
        sv.event_record(1)
        wall.append((time.perf_counter() - t0) * 1e3)
        dev.append(sv.event_elapsed(0, 1))
    m = len(uniforms)
    ms = min(dev)
    b = S + 16 * m
    out["samples"] = {"count": m, "ms": round(ms, 2), "wall_ms": round(min(wall), 2),
                      "algorithmic_GB": round(b / 1e9, 2), "GBs": round(b / (ms / 1e3) / 1e9, 1),
                      "hbm_frac": round(b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                      "first_indices": [int(i) for i in idx[:4]]}
    log(f"<H> = {e:.12f}: {out['expval_100_terms']['ms']} ms ({out['expval_100_terms']['GBs']} GB/s); "
        f"{m} samples: {out['samples']['ms']} ms ({out['samples']['GBs']} GB/s)")
    # marginal probabilities of up to 8 scattered wires (pl_probs): one read of the
    # state, 2^8 bins back to the host
    wires = [w for w in (0, 3, 7, 11, 15, 19, 23) if w < n - 1] + [n - 1]
    p = sv.probs(wires)
    dev = []
    for _ in range(reps):
        sv.synchronize()
        sv.event_record(0)
        p = sv.probs(wires)
        sv.event_record(1)
        dev.append(sv.event_elapsed(0, 1))
    ms = min(dev)
    out["probs_8_wires"] = {"wires": wires, "ms": round(ms, 3), "sum": float(p.sum()),
                            "GBs": round(S / (ms / 1e3) / 1e9, 1),
                            "hbm_frac": round(S / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"probs of {len(wires)} wires: {out['probs_8_wires']['ms']} ms ({out['probs_8_wires']['GBs']} GB/s)")
    return out


def bench_precision(n_big=33, n_ref=30):
    """Precision form: the same seeded 50-layer gate mix fused into 5-qubit
    32x32 blocks (operations move ahead only past operations on disjoint
    wires) applied to a complex64 state on FP32 matrix cores, renormalised in
    FP64 after every block; and the same circuit as complex64 tile passes,
    renormalised in FP64 at the end.
    Error reference at n_ref qubits: the fused complex64 run against the
    unfused complex128 run of the same circuit (|d<H>|, ||d psi||_2, norm
    drift without renormalisation).  Throughput at n_big qubits (64 GiB)."""
    from lightning_unification_amd import capi
    out = {}
    # error reference
    tape = random_layers(n_ref, 50, seed=1002)
    ham = random_hamiltonian(n_ref, 100, seed=1002)
    s128 = capi.DeviceStateVector(n_ref, np.complex128)
    s128.apply_tape(tape)
    e128 = s128.expval(ham)
    s64 = capi.DeviceStateVector(n_ref, np.complex64)
    nb = s64.apply_tape_fused(tape, 5, renorm=True)
    e64 = s64.expval(ham)
    dist = s64.distance(s128)
    s64.reset()
    s64.apply_tape_fused(tape, 5, renorm=False)
    drift = abs(s64.norm2() - 1.0)
    s64.close()
    s64 = capi.DeviceStateVector(n_ref, np.complex64)
    s64.apply_tape_tiled(tape)
    drift_t = abs(s64.norm2() - 1.0)
    s64.reset()
    s64.apply_tape_tiled(tape, renorm=True)
    et = s64.expval(ham)
    dist_t = s64.distance(s128)
    s64.close()
    s128.close()
    out["error_vs_c128"] = {"qubits": n_ref, "blocks": nb, "abs_dH": abs(e64 - e128),
                            "H_c128": e128, "state_l2": dist,
                            "norm_drift_without_renorm": drift,
                            "tiled_c64": {"abs_dH": abs(et - e128), "state_l2": dist_t,
                                          "norm_drift_without_renorm": drift_t}}
    log(f"precision n={n_ref}: |dH| {abs(e64 - e128):.2e}, |dpsi| {dist:.2e}, drift {drift:.2e}")

    # throughput at n_big
    warm = capi.DeviceStateVector(14, np.complex64)
    warm.apply_tape_fused(random_layers(14, 2, seed=1), 5, renorm=True)
    warm.close()
    tape = random_layers(n_big, 50, seed=1002)
    ham = random_hamiltonian(n_big, 100, seed=1002)
    sv = capi.DeviceStateVector(n_big, np.complex64)
    sv.synchronize()
    sv.timer_start()
    nb = sv.apply_tape_fused(tape, 5, renorm=True)
    ms = sv.timer_stop()
    S = (1 << n_big) * 8
    gate_bytes = sum(2 * touched_fraction(op[0]) * S for op in tape.ops)
    block_bytes = nb * 2 * S  # every block reads and writes the state once
    # a 32x32 complex block per group of 32 amplitudes: 8 * 32 * 32 FLOP of
    # complex arithmetic, executed as four real 32x32 products (the real
    # 64x64 form [[Ur, -Ui], [Ui, Ur]]), so the matrix cores execute as many
    flops = nb * (1 << (n_big - 5)) * 8 * 32 * 32
    mfma_flops = nb * (1 << (n_big - 5)) * 4 * 2 * 32 * 32
    t0 = time.perf_counter()
    e = sv.expval(ham)
    t_h = time.perf_counter() - t0
    norm = sv.norm2()
    sv.reset()
    sv.synchronize()
    sv.timer_start()
    npass = sv.apply_tape_tiled(tape, renorm=True)
    ms_t = sv.timer_stop()
    norm_t = sv.norm2()
    sv.close()
    out["c64_tiled"] = {"qubits": n_big, "gates": len(tape), "passes": npass, "ms": round(ms_t, 1),
                        "hbm_GBs": round(npass * 2 * S / (ms_t / 1e3) / 1e9, 1),
                        "gate_equivalent_GBs": round(gate_bytes / (ms_t / 1e3) / 1e9, 1),
                        "norm": norm_t}
    log(f"precision n={n_big} tiled: {npass} passes in {ms_t:.0f} ms")
    out["c64_fused"] = {"qubits": n_big, "gates": len(tape), "blocks": nb, "ms": round(ms, 1),
                        "block_GBs": round(block_bytes / (ms / 1e3) / 1e9, 1),
                        "gate_equivalent_GBs": round(gate_bytes / (ms / 1e3) / 1e9, 1),
                        "complex_TFLOPs": round(flops / (ms / 1e3) / 1e12, 2),
                        "mfma_TFLOPs": round(mfma_flops / (ms / 1e3) / 1e12, 2),
                        "mfma_frac": round(mfma_flops / (ms / 1e3) / 1e12 / 157.3, 4),
                        "hbm_frac": round(block_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                        "H": e, "H_ms": round(t_h * 1e3, 1), "norm": norm}
    log(f"precision n={n_big}: {nb} blocks in {ms:.0f} ms, "
        f"{block_bytes / (ms / 1e3) / 1e9:.0f} GB/s, {mfma_flops / (ms / 1e3) / 1e12:.1f} TFLOP/s")
    return out


def bench_surrogate(samples=1 << 20, epochs=3):
    """Learned-surrogate form: 16-qubit, 6-layer HEA (192 angles), periodic
    TFIM (J = h = 1).  `samples` uniform (theta, E, grad E) tuples from the
    batched adjoint engine (complex64), an MLP 192 -> 512 -> 512 -> 193
    trained on them with the BF16 matrix-core kernels, then: held-out R^2,
    the screening gain (network's pick among 256 random candidates vs a random
    pick, scored exactly), and gradient descent with and without screening at
    equal exact-call budgets.  The same measures on the one-layer ansatz show
    the method where the landscape is learnable."""
    from lightning_unification_amd import surrogate as S
    out = {}
    for tag, depth, kind, nsamp, ep in (("hea16x6", 6, "angle", samples, epochs),
                                        ("hea16x1", 1, "periodic", samples // 4, 6)):
        t0 = time.perf_counter()
        prob = S.Problem(16, depth, batch=4096, dtype=np.complex64)
        th, E, G = S.generate(prob, nsamp, seed=1)
        tht, Et, Gt = S.generate(prob, 8192, seed=2)
        prob.close()
        t_gen = time.perf_counter() - t0
        sur = S.Surrogate(prob.nparams, kind=kind)
        t0 = time.perf_counter()
        losses = sur.fit(th, E, G, epochs=ep, batch=1024, lr=1e-3)
        t_fit = time.perf_counter() - t0
        d = sur.net.dims
        flops = 6.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[3]) * (nsamp // 1024 * 1024) * ep
        e_hat, g_hat = sur.predict(tht)
        exact = S.Problem(16, depth, batch=256)
        gain = S.screen_gain(exact, sur, rounds=8)
        th0 = S.sample_thetas(1, prob.nparams, 100)[0]
        h0, c0, _ = S.optimize(exact, th0, target=-1e9, max_calls=100, eta=0.05)
        h1, c1, _ = S.optimize(exact, th0, target=-1e9, max_calls=100, eta=0.05, surrogate=sur)
        exact.close()
        sur.close()
        out[tag] = {"layers": depth, "params": prob.nparams, "features": kind, "net": list(d),
                    "tuples": nsamp, "gen_s": round(t_gen, 1),
                    "tuples_per_s": round((nsamp + 8192) / t_gen, 1),
                    "epochs": ep, "train_s": round(t_fit, 2),
                    "train_TFLOPs": round(flops / t_fit / 1e12, 2),
                    "loss_first_last": [round(losses[0], 4), round(losses[-1], 4)],
                    "heldout_R2_E": round(S.r2(e_hat, Et), 4), "heldout_R2_grad": round(S.r2(g_hat, Gt), 4),
                    "screen_256": gain,
                    "gd_100_exact_calls": {"plain_E": h0[-1], "screened_E": h1[-1],
                                           "plain_calls": c0, "screened_calls": c1}}
        log(f"surrogate {tag}: R2 E {out[tag]['heldout_R2_E']}, screen {gain}, "
            f"GD plain {h0[-1]:.4f} vs screened {h1[-1]:.4f}")
    out["hea16x6_trajectory"] = bench_surrogate_trajectory()
    return out


def bench_surrogate_trajectory(budget=120):
    """The learned-surrogate form at its stated size (16 qubits, 6 layers,
    192 angles, 192 -> 512 -> 512 -> 193 on BF16 matrix cores) with the
    network trained around the optimiser's trajectory: every 8 accepted
    steps 4096 points theta + 0.05 N(0, 1) are evaluated by the batched
    adjoint engine (complex128) and the network is retrained on them; each
    step it screens 256 step lengths along -grad E and the exact simulation
    verifies its pick.  Compared with plain gradient descent (same start,
    eta = 0.05, halved on rejection) at equal budgets of exact calls inside
    the optimisation loop; the training evaluations are reported apart."""
    from lightning_unification_amd import surrogate as S
    exact = S.Problem(16, 6, batch=4096)
    th0 = S.sample_thetas(1, exact.nparams, 100)[0]
    t0 = time.perf_counter()
    tr_plain = []
    S.optimize(exact, th0, target=-1e9, max_calls=budget, eta=0.05, trace=tr_plain)
    t_plain = time.perf_counter() - t0
    sur = S.Surrogate(exact.nparams, kind="local", seed=3)
    t0 = time.perf_counter()
    tr_loc = []
    _, loop_calls, data_calls, _ = S.optimize_local(exact, th0, target=-1e9, max_calls=budget, surrogate=sur,
                                                    trace=tr_loc)
    t_loc = time.perf_counter() - t0
    sur.close()
    exact.close()
    e_plain = tr_plain[-1][1]
    out = {"budget_loop_calls": budget, "start_E": tr_plain[0][1],
           "plain": {"E": e_plain, "calls": tr_plain[-1][0], "accepted": len(tr_plain), "s": round(t_plain, 2)},
           "screened": {"E": tr_loc[-1][1], "loop_calls": loop_calls, "training_evaluations": data_calls,
                        "accepted": len(tr_loc), "s": round(t_loc, 2),
                        "loop_calls_to_reach_plain_final_E": S.calls_to_reach(tr_loc, e_plain)},
           "E_at_calls": {str(c): {"plain": S.energy_at(tr_plain, c), "screened": S.energy_at(tr_loc, c)}
                          for c in (10, 20, 40, 80, budget)}}
    log(f"surrogate trajectory: plain {e_plain:.4f} vs screened {tr_loc[-1][1]:.4f} at {budget} loop calls "
        f"({data_calls} training evaluations)")
    return out


def bench_stream(n=30, reps=10):
    """Streaming ceilings of the device for complex128 amplitudes at the
    headline state size (16 GiB): read only, write only, in-place
    read-modify-write, and the pair pattern of a dense 1-qubit gate on target
    bit p (the access pattern of k_group_flat<double, 2>), with the engine's
    non-temporal loads and stores (csrc/stream_probe.hip).  GB/s count the
    bytes each kernel must move: 16 B per amplitude read and/or written."""
    import ctypes
    from lightning_unification_amd import capi
    lib = ctypes.CDLL(os.path.join(ROOT, "lightning_unification_amd", "libpl_stream.so"))
    lib.sp_run.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    lib.sp_last_error.restype = ctypes.c_char_p
    sv = capi.DeviceStateVector(n, np.complex128)
    scratch = capi.DeviceStateVector(max(1, n - 8 - 1), np.complex128)  # >= one double per workgroup
    dev, sdev = sv.device_ptr(), scratch.device_ptr()
    amps = 1 << n
    out = {"qubits": n, "state_GiB": amps * 16 / 2 ** 30, "reps": reps}

    def run(mode, p=0):
        ms = ctypes.c_double()
        lib.sp_run(ctypes.c_void_p(dev), amps, mode, p, 1, ctypes.c_void_p(sdev), ctypes.byref(ms))  # warm
        rc = lib.sp_run(ctypes.c_void_p(dev), amps, mode, p, reps, ctypes.c_void_p(sdev), ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(lib.sp_last_error().decode())
        return ms.value

    for name, mode, mult in (("read", 0, 1), ("write", 1, 1), ("rmw", 2, 2)):
        ms = run(mode)
        out[name] = {"ms": round(ms, 4), "GBs": round(mult * amps * 16 / (ms / 1e3) / 1e9, 1)}
        log(f"stream {name}: {ms:.3f} ms, {out[name]['GBs']} GB/s")
    pairs = {}
    for p in (0, 3, 6, 8, 12, 16, 20, 25, 29):
        if p < n:
            ms = run(3, p)
            pairs[str(p)] = {"ms": round(ms, 4), "GBs": round(2 * amps * 16 / (ms / 1e3) / 1e9, 1)}
            log(f"stream pair p={p}: {ms:.3f} ms, {pairs[str(p)]['GBs']} GB/s")
    out["rmw_pair_by_target_bit"] = pairs
    sv.close()
    scratch.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--shots", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=4096,
                    help="circuits of the batching workload (0 skips it)")
    ap.add_argument("--plain-s", type=float, default=8.0,
                    help="seconds of the per-circuit path timed for the batching comparison")
    ap.add_argument("--precision", type=int, default=33,
                    help="qubits of the complex64 fused-block run (0 skips the precision form)")
    ap.add_argument("--surrogate", type=int, default=1 << 20,
                    help="training tuples of the learned-surrogate form (0 skips it)")
    ap.add_argument("--no-c1", action="store_true", help="skip the 20-qubit adjoint measurement")
    ap.add_argument("--no-tiled", action="store_true", help="skip the tile-pass measurement")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the VJP, sparse-H and Metropolis measurements")
    ap.add_argument("--mode", choices=["gates", "copy", "stream", "measure", "batch", "precision", "surrogate",
                                       "c1", "next"],
                    default="gates",
                    help="copy: device-to-device copies of a 4 GiB state, the plain copy "
                         "kernel used to scale rocprof FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the sharded "
              "multi-GPU path is a design, not part of this build.", file=sys.stderr)
        sys.exit(2)

    from lightning_unification_amd import capi
    if args.mode == "copy":
        a = capi.DeviceStateVector(28, np.complex128)
        b = capi.DeviceStateVector(28, np.complex128)
        for _ in range(3):
            b.copy_from(a)
        b.synchronize()
        print(json.dumps({"mode": "copy", "bytes_per_copy": (1 << 28) * 16, "copies": 3}))
        return
    if args.mode == "measure":
        # the benchmark circuit's final state (tile passes), then the
        # measurements alone: the run to profile k_pauli_tile and the sampler
        n = args.qubits
        sv = capi.DeviceStateVector(n, np.complex128)
        sv.apply_tape_tiled(random_layers(n, 50, seed=1002))
        res = bench_measurements(sv, random_hamiltonian(n, 100, seed=1002),
                                 sample_uniforms(args.shots, seed=1002), n)
        sv.close()
        print(json.dumps({"mode": "measure", **res}))
        return
    if args.mode == "stream":
        print(json.dumps({"mode": "stream", "stream": bench_stream(args.qubits)}))
        return
    if args.mode == "batch":
        print(json.dumps({"mode": "batch", "batched": bench_batch(args.batch, args.plain_s)}))
        return
    if args.mode == "precision":
        print(json.dumps({"mode": "precision", "precision": bench_precision(args.precision)}))
        return
    if args.mode == "c1":
        print(json.dumps({"mode": "c1", "c1": bench_c1()}))
        return
    if args.mode == "next":
        print(json.dumps({"mode": "next", "next_rows": bench_next()}))
        return
    if args.mode == "surrogate":
        print(json.dumps({"mode": "surrogate", "surrogate": bench_surrogate(args.surrogate)}))
        return
    n = args.qubits
    tape = random_layers(n, 50, seed=1002)
    ham = random_hamiltonian(n, 100, seed=1002)
    uniforms = sample_uniforms(args.shots, seed=1002)
    gate_bytes = [2 * touched_fraction(op[0]) * (1 << n) * 16 for op in tape.ops]
    log(f"workload: {len(tape)} gates on {n} qubits, {sum(gate_bytes)/1e9:.1f} GB per pass")

    sv = capi.DeviceStateVector(n, np.complex128)
    for _ in range(args.warmup):
        sv.apply_tape(tape)
    sv.synchronize()
    log("warmup done")

    # timed region: events around every launch on the engine's stream
    per_gate_ms = np.zeros(len(tape))
    sv.synchronize()
    sv.event_record(0)
    slot = 1
    for _ in range(args.steps):
        for op in tape.ops:
            sv.apply(op[0], op[1], op[2], op[3])
            sv.event_record(slot)
            slot += 1
    total_ms = sv.event_elapsed(0, slot - 1)
    s = 1
    for step in range(args.steps):
        prev = 0 if step == 0 else s - 1
        for i in range(len(tape)):
            per_gate_ms[i] += sv.event_elapsed(prev, s)
            prev = s
            s += 1
    sv.synchronize()
    log(f"timed {args.steps} steps: {total_ms:.1f} ms")

    total_bytes = sum(gate_bytes) * args.steps
    value = total_bytes / (total_ms / 1e3) / 1e9

    # dominant kernel: sum time and algorithmic bytes per kernel instantiation
    by_kernel = {}
    for i, op in enumerate(tape.ops):
        k = kernel_for(op[0], op[1], n)
        d = by_kernel.setdefault(k, {"ms": 0.0, "bytes": 0.0, "launches": 0, "gates": set()})
        d["ms"] += per_gate_ms[i]
        d["bytes"] += gate_bytes[i] * args.steps
        d["launches"] += args.steps
        d["gates"].add(op[0])
    dom = max(by_kernel, key=lambda k: by_kernel[k]["ms"])
    dk = by_kernel[dom]
    traffic, traffic_src = profiled_traffic(dom) if n == 30 else (None, None)
    achieved = dk["bytes"] / (dk["ms"] / 1e3) / 1e9
    per_gate_kind = {}
    for i, op in enumerate(tape.ops):
        d = per_gate_kind.setdefault(op[0], [0.0, 0.0])
        d[0] += gate_bytes[i] * args.steps
        d[1] += per_gate_ms[i]
    per_gate_gbs = {g: round(b / (ms / 1e3) / 1e9, 1) for g, (b, ms) in per_gate_kind.items()}

    # measurements on the final state
    meas = bench_measurements(sv, ham, uniforms, n)
    tiled = None if args.no_tiled else bench_tiled(sv, tape, n, args.warmup + args.steps)
    sv.close()

    c1 = None if args.no_c1 else bench_c1()
    next_rows = None if args.no_next else bench_next()
    batched = bench_batch(args.batch, args.plain_s) if args.batch > 0 else None
    precision = bench_precision(args.precision) if args.precision > 0 else None
    surrogate = bench_surrogate(args.surrogate) if args.surrogate > 0 else None
    cpu = None if args.no_cpu else cpu_baseline(tape, n)

    out = {
        "metric": "HBM GB/s gate-apply @30q c128; batched adjoint grads/s at 1e-6 err vs plain",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded random circuit, SplitMix64)",
        "config": {"workload": f"random 50-layer circuit, {n} qubits, complex128 state in HBM, "
                               f"{len(tape)} gates per step", "n_qubits": n, "gates": len(tape),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "launches": dk["launches"],
                     "avg_launch_ms": round(dk["ms"] / dk["launches"], 4),
                     "algorithmic_bytes_per_launch": round(dk["bytes"] / dk["launches"]),
                     "traffic": traffic, "traffic_source": traffic_src},
        "cpu_baseline": cpu,
        "per_gate_GBs": per_gate_gbs,
        "expval_100_terms": meas["expval_100_terms"],
        "samples": meas["samples"],
        "probs_8_wires": meas["probs_8_wires"],
        "tiled_circuit": tiled,
        "c1_adjoint": c1,
        "next_rows": next_rows,
        "batched": batched,
        "precision": precision,
        "surrogate": surrogate,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
