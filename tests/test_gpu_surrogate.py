"""Learned surrogate end to end on the GPU: tuples from the batched adjoint
engine, training with the BF16 matrix-core MLP, screening with exact
verification.  The gain is shown on a landscape the network can learn (one
HEA layer); every accepted optimiser step is checked against a fresh exact
evaluation."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    from lightning_unification_amd import surrogate as S
    prob = S.Problem(12, 1, batch=2048, dtype=np.complex64)
    th, E, G = S.generate(prob, 65536, seed=11)
    prob.close()
    sur = S.Surrogate(24, hidden=(256, 256), kind="periodic", seed=5)
    losses = sur.fit(th, E, G, epochs=8, batch=512, lr=1e-3)
    exact = S.Problem(12, 1, batch=256)
    yield S, sur, exact, losses
    sur.close()
    exact.close()


def test_surrogate_learns_and_screens(trained):
    S, sur, exact, losses = trained
    assert losses[-1] < 0.9 * losses[0]
    th = S.sample_thetas(2048, 24, 99)
    E, G = exact.exact(th)
    e_hat, g_hat = sur.predict(th)
    assert S.r2(e_hat, E) > 0.6
    gain = S.screen_gain(exact, sur, rounds=12)
    print("screen gain", gain, flush=True)
    # the network's pick beats a random pick: lower mean exact energy, and
    # below the candidates' median in most rounds
    assert gain["pick"] < gain["random"] - 0.5 * (gain["random"] - gain["best"])
    assert gain["win_rate"] >= 0.75


def test_screened_optimiser_is_exact(trained):
    S, sur, exact, _ = trained
    th0 = S.sample_thetas(1, 24, 7)[0]
    hist, calls, th = S.optimize(exact, th0, target=-1e9, max_calls=40, eta=0.05, surrogate=sur)
    assert calls <= 41 and len(hist) >= 2
    assert np.all(np.diff(hist) < 0)                  # every accepted step lowers E
    (e_final,), _ = exact.exact(th[None])             # and its energy is the exact one
    assert abs(e_final - hist[-1]) < 1e-12 * max(1.0, abs(e_final))
