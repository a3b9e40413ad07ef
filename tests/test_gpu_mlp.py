"""The surrogate's MLP kernels (pl_mlp_*, BF16 matrix cores, FP32
accumulation) against a plain PyTorch fp32 reference of the same network on
the CPU: outputs, loss and every parameter gradient, then Adam training on a
known smooth function."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (192, 512, 512, 193)


def bf16(a):
    import torch
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float32)


def torch_reference(net, params, x, t, w):
    """Forward + backward in fp32 on the CPU with the operands the device
    rounds to BF16 (features, weights, hidden activations) rounded alike."""
    import torch
    layers = net.unpack(params)
    W = [bf16(Wl).requires_grad_(True) for Wl, _ in layers]
    b = [torch.tensor(bl, requires_grad=True) for _, bl in layers]
    a = bf16(x)
    for l in range(2):
        z = a @ W[l] + b[l]
        g = torch.nn.functional.gelu(z)                # exact (erf) GELU
        a = g + (g.to(torch.bfloat16).to(torch.float32) - g).detach()  # BF16 activation, straight-through
    y = a @ W[2] + b[2]
    loss = ((y - torch.tensor(t)) ** 2 * torch.tensor(w)).sum() / x.shape[0]
    loss.backward()
    grads = []
    for l in range(3):
        grads += [W[l].grad.numpy().reshape(-1), b[l].grad.numpy()]
    return y.detach().numpy(), float(loss.detach()), np.concatenate(grads)


def test_forward_and_gradients_match_torch():
    from lightning_unification_amd import capi
    rs = np.random.default_rng(0)
    net = capi.Mlp(DIMS, max_batch=128, seed=7)
    p = net.params()
    n = 96
    x = rs.uniform(-1, 1, (n, DIMS[0])).astype(np.float32)
    t = rs.normal(size=(n, DIMS[3])).astype(np.float32)
    w = rs.uniform(0.5, 1.5, DIMS[3]).astype(np.float32)
    net.set_loss_weights(w)
    y_ref, loss_ref, g_ref = torch_reference(net, p, x, t, w)
    y = net.predict(x)
    assert np.max(np.abs(y - y_ref)) < 2e-3 * np.max(np.abs(y_ref))
    loss, g = net.gradients(x, t)
    assert abs(loss - loss_ref) < 1e-3 * loss_ref
    # per tensor: the device also rounds the output and hidden gradients to
    # BF16 inside its products, so gradients agree to BF16 precision
    o = 0
    for l in range(3):
        for size in (DIMS[l] * DIMS[l + 1], DIMS[l + 1]):
            a, r = g[o:o + size], g_ref[o:o + size]
            o += size
            assert np.max(np.abs(a - r)) < 3e-2 * np.max(np.abs(r)), (l, size)
            cos = float(np.dot(a, r) / (np.linalg.norm(a) * np.linalg.norm(r)))
            assert cos > 0.999, (l, size, cos)
    net.close()


def test_training_fits_a_smooth_function():
    from lightning_unification_amd import capi
    rs = np.random.default_rng(1)
    dims = (16, 128, 128, 3)
    net = capi.Mlp(dims, max_batch=256, seed=3)
    A = rs.normal(size=(16, 3)) / 4
    x = rs.uniform(-1, 1, (8192, 16)).astype(np.float32)
    t = np.column_stack([np.sin(x @ A[:, 0]), np.cos(x @ A[:, 1]), (x @ A[:, 2]) ** 2]).astype(np.float32)
    net.load_data(x, t)
    first = net.train(32, 256, 3e-3)
    for _ in range(15):
        last = net.train(32, 256, 3e-3)
    assert last < 0.05 * first
    xt = rs.uniform(-1, 1, (512, 16)).astype(np.float32)
    tt = np.column_stack([np.sin(xt @ A[:, 0]), np.cos(xt @ A[:, 1]), (xt @ A[:, 2]) ** 2])
    assert np.sqrt(np.mean((net.predict(xt) - tt) ** 2)) < 0.05
    # parameters round-trip; set_params resets the optimiser state
    p = net.params()
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    net.close()


def test_mlp_rejects_bad_batch():
    from lightning_unification_amd import capi
    net = capi.Mlp((4, 8, 8, 2), max_batch=16)
    with pytest.raises(ValueError):
        net.gradients(np.zeros((17, 4)), np.zeros((17, 2)))
    with pytest.raises(ValueError):
        net.train(1, 8, 1e-3)  # no data loaded
    net.close()
