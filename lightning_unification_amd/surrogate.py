"""Learned surrogate for a layered-ansatz energy landscape.

Workload: an n-qubit hardware-efficient ansatz (per layer RY, RZ on every
wire, then a CNOT ring) with a periodic transverse-field Ising Hamiltonian.
The batched adjoint engine (pl_batch_*) generates (theta, E, grad E) tuples;
an MLP (pl_mlp_*: BF16 matrix cores, FP32 accumulation) learns
theta -> (E, grad E); an optimiser uses it to screen candidate steps, and the
exact adjoint simulation verifies every step it accepts.

Features: "angle" gives x = theta / pi - 1 in [-1, 1) per angle (P inputs, the
192 -> 512 -> 512 -> 193 network of the 16-qubit, 6-layer workload);
"periodic" gives (cos theta, sin theta) per angle (2P inputs): E is affine in
(cos theta_i, sin theta_i) for every single angle, so this input carries the
landscape's periodicity.  Targets: E and every gradient component, each
standardised by the training set's mean and standard deviation.
"""
import math

import numpy as np

from . import capi
from .circuits import Rng, tfim


class Problem:
    """Exact energies and gradients of the ansatz, many circuits per launch."""

    def __init__(self, n=16, depth=6, h=1.0, batch=4096, dtype=np.complex128):
        self.n, self.depth, self.batch = n, depth, batch
        self.nparams = 2 * n * depth
        self.ham = tfim(n, 1.0, h)
        self.plans = {b: capi.BatchPlan([n] * b, [depth] * b, "YZ", "ring", [self.ham] * b, dtype=dtype)
                      for b in sorted({1, batch})}
        self.calls = 0  # exact (E, grad E) circuit evaluations requested

    def exact(self, thetas):
        """-> E[k], G[k, P] for thetas[k, P] (any k; padded to the batch)."""
        thetas = np.atleast_2d(np.asarray(thetas, np.float64))
        k = thetas.shape[0]
        self.calls += k
        E = np.empty(k)
        G = np.empty((k, self.nparams))
        b = 1 if k == 1 else self.batch
        plan = self.plans[b]
        for r in range(0, k, b):
            chunk = thetas[r:r + b]
            m = chunk.shape[0]
            p = np.zeros((b, self.nparams))
            p[:m] = chunk
            e, g = plan.run(p.reshape(-1))
            E[r:r + m] = e[:m]
            G[r:r + m] = g.reshape(b, self.nparams)[:m]
        return E, G

    def close(self):
        for p in self.plans.values():
            p.close()


def features(theta, kind="angle", center=None, scale=1.0):
    theta = np.asarray(theta, np.float64)
    if kind == "periodic":
        return np.concatenate([np.cos(theta), np.sin(theta)], axis=-1).astype(np.float32)
    if kind == "local":  # offsets from a trust-region centre, in units of its radius
        return ((theta - center) / scale).astype(np.float32)
    return (theta / math.pi - 1.0).astype(np.float32)


def sample_thetas(count, nparams, seed):
    return 2.0 * math.pi * Rng(seed).uniform(count * nparams).reshape(count, nparams)


def generate(problem, count, seed=1005):
    """count uniform (theta, E, grad E) tuples from the batched engine."""
    th = sample_thetas(count, problem.nparams, seed)
    E, G = problem.exact(th)
    return th, E, G


class Surrogate:
    """MLP theta -> (E, grad E) with standardised targets."""

    def __init__(self, nparams, hidden=(512, 512), max_batch=1024, seed=1005, kind="angle"):
        self.nparams = nparams
        self.kind = kind
        d0 = 2 * nparams if kind == "periodic" else nparams
        self.net = capi.Mlp((d0, hidden[0], hidden[1], nparams + 1), max_batch, seed)
        self.mu = np.zeros(nparams + 1)
        self.sd = np.ones(nparams + 1)
        self.center = np.zeros(nparams)
        self.scale = 1.0

    def targets(self, E, G):
        return ((np.column_stack([E, G]) - self.mu) / self.sd).astype(np.float32)

    def fit(self, theta, E, G, epochs=4, batch=1024, lr=1e-3, energy_weight=None):
        """Train on the tuples (resident on the device); returns per-epoch mean loss."""
        T = np.column_stack([E, G])
        self.mu = T.mean(axis=0)
        self.sd = T.std(axis=0) + 1e-12
        w = np.ones(self.nparams + 1, np.float32)
        # the energy is one output of nparams + 1: weight it like all gradients together
        w[0] = float(self.nparams) if energy_weight is None else energy_weight
        self.net.set_loss_weights(w / w.sum())
        self.net.load_data(features(theta, self.kind, self.center, self.scale), self.targets(E, G))
        steps = max(1, theta.shape[0] // batch)
        return [self.net.train(steps, batch, lr) for _ in range(epochs)]

    def predict(self, theta):
        y = self.net.predict(features(np.atleast_2d(theta), self.kind, self.center, self.scale)).astype(np.float64)
        y = y * self.sd + self.mu
        return y[:, 0], y[:, 1:]

    def close(self):
        self.net.close()


def candidate_steps(eta, count=256, span=4.0):
    """Step sizes eta * 2^u, u evenly spaced in [-span, span]."""
    return eta * np.exp2(np.linspace(-span, span, count))


def optimize(problem, theta0, target, max_calls=400, eta=0.05, surrogate=None, candidates=256, trace=None):
    """Gradient descent from theta0 until E <= target or max_calls exact calls.

    Without a surrogate: theta - eta * g, verified exactly; a rejected step
    halves eta.  With one: the surrogate predicts E for `candidates` step
    sizes along -g and the exact simulation verifies the best; a rejected
    pick falls back to the plain step.  Every accepted step is exact.
    Returns (E history of accepted steps, exact calls used, final theta);
    `trace`, if a list, receives (exact calls so far, E) per accepted step."""
    start_calls = problem.calls
    theta = np.array(theta0, np.float64)
    (E,), (g,) = problem.exact(theta[None])
    hist = [E]
    if trace is not None:
        trace.append((1, E))
    while E > target and problem.calls - start_calls < max_calls:
        tried = False
        if surrogate is not None:
            steps = candidate_steps(eta, candidates)
            cand = theta[None, :] - steps[:, None] * g[None, :]
            e_hat, _ = surrogate.predict(cand)
            pick = int(np.argmin(e_hat))
            (E1,), (g1,) = problem.exact(cand[pick][None])
            tried = True
            if E1 < E:
                theta, E, g = cand[pick], E1, g1
                eta = float(steps[pick])  # the plain fallback continues from the screened step
                hist.append(E)
                if trace is not None:
                    trace.append((problem.calls - start_calls, E))
                continue
        while problem.calls - start_calls < max_calls:
            cand = theta - eta * g
            (E1,), (g1,) = problem.exact(cand[None])
            if E1 < E:
                theta, E, g = cand, E1, g1
                hist.append(E)
                if trace is not None:
                    trace.append((problem.calls - start_calls, E))
                break
            eta *= 0.5
            if tried:
                break
    return np.array(hist), problem.calls - start_calls, theta


def optimize_local(problem, theta0, target, max_calls, surrogate, sigma=0.05, samples=4096, refresh=8,
                   epochs=6, batch=512, lr=1e-3, candidates=256, eta=0.05, seed=77, trace=None):
    """Gradient descent screened by a surrogate trained around the iterate.

    Every `refresh` accepted steps the batched engine evaluates `samples`
    points theta + sigma * N(0, 1) around the current iterate (the training
    evaluations, counted apart) and the network, in features centred there,
    is retrained on them.  Each step it predicts E for `candidates` points
    theta - s * g / |g| with s from sigma / 16 to 4 sigma, and the exact
    simulation verifies the best; a rejected pick falls back to the plain
    step (theta - eta g, eta halved on rejection).  Every accepted step is
    exact.  Returns (E history of accepted steps, loop exact calls, training
    evaluations, final theta); `trace`, if a list, receives (loop exact calls
    so far, E) per accepted step."""
    start_calls = problem.calls
    theta = np.array(theta0, np.float64)
    (E,), (g,) = problem.exact(theta[None])
    loop_calls = 1
    data_calls = 0
    hist = [E]
    if trace is not None:
        trace.append((1, E))
    rs = np.random.default_rng(seed)
    since = refresh
    while E > target and loop_calls < max_calls:
        if since >= refresh:
            th = theta[None, :] + sigma * rs.standard_normal((samples, theta.size))
            Ed, Gd = problem.exact(th)
            data_calls += samples
            surrogate.center = theta.copy()
            surrogate.scale = sigma
            surrogate.fit(th, Ed, Gd, epochs=epochs, batch=batch, lr=lr)
            since = 0
        gn = float(np.linalg.norm(g))
        steps = sigma * np.exp2(np.linspace(-4.0, 2.0, candidates))
        cand = theta[None, :] - (steps / max(gn, 1e-300))[:, None] * g[None, :]
        e_hat, _ = surrogate.predict(cand)
        pick = int(np.argmin(e_hat))
        (E1,), (g1,) = problem.exact(cand[pick][None])
        loop_calls += 1
        if E1 < E:
            theta, E, g = cand[pick], E1, g1
            hist.append(E)
            if trace is not None:
                trace.append((loop_calls, E))
            since += 1
            continue
        while loop_calls < max_calls:
            cand1 = theta - eta * g
            (E1,), (g1,) = problem.exact(cand1[None])
            loop_calls += 1
            if E1 < E:
                theta, E, g = cand1, E1, g1
                hist.append(E)
                if trace is not None:
                    trace.append((loop_calls, E))
                since += 1
                break
            eta *= 0.5
        since = refresh  # the model missed: retrain around the new iterate
    assert problem.calls - start_calls == loop_calls + data_calls
    return np.array(hist), loop_calls, data_calls, theta


def energy_at(trace, calls):
    """Energy of the last accepted step within `calls` exact calls."""
    e = None
    for c, E in trace:
        if c <= calls:
            e = E
    return e


def calls_to_reach(trace, target):
    """Exact calls until the first accepted step at or below `target`."""
    for c, E in trace:
        if E <= target:
            return c
    return None


def r2(pred, true):
    """Coefficient of determination over all entries."""
    pred, true = np.asarray(pred), np.asarray(true)
    return float(1.0 - np.mean((pred - true) ** 2) / np.var(true))


def screen_gain(problem, surrogate, rounds=16, candidates=256, seed=2024):
    """Screening value of the network: each round draws `candidates` uniform
    parameter vectors, the network picks the one with the lowest predicted
    energy, and the exact simulation scores every candidate (one batched
    launch).  Returns mean exact energy of the network's pick, of a random
    pick (the candidates' mean), of the exact best, and the fraction of
    rounds in which the pick beats the candidates' median."""
    pick_e, mean_e, best_e, wins = [], [], [], 0
    for r in range(rounds):
        cand = sample_thetas(candidates, problem.nparams, seed + r)
        e_hat, _ = surrogate.predict(cand)
        e, _ = problem.exact(cand)
        k = int(np.argmin(e_hat))
        pick_e.append(e[k])
        mean_e.append(e.mean())
        best_e.append(e.min())
        wins += e[k] < np.median(e)
    return {"pick": float(np.mean(pick_e)), "random": float(np.mean(mean_e)),
            "best": float(np.mean(best_e)), "win_rate": wins / rounds}
