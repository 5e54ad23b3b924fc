"""The learned regrid predictor's HIP kernels (amrlbm_cnn.hip) against a plain
PyTorch FP32 reference of the same network on the CPU. The reference rounds
the input, the weights and the activations between layers to BF16 as the
kernels store them (rounding passed straight through in the backward pass),
so forward logits and weight gradients agree to FP32 summation order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
F = torch.nn.functional


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def _unpack(w):
    """Flat parameter vector -> torch conv tensors [cout, cin, 3, 3] and biases."""
    w = torch.tensor(w, dtype=torch.float32)
    out, o = [], 0
    for cout, cin in ((16, 3), (16, 16), (2, 16)):
        flat = w[o:o + cout * cin * 9].reshape(cout, 9, cin)          # [cout][tap][cin]
        out.append(flat.permute(0, 2, 1).reshape(cout, cin, 3, 3))    # tap = 3 dy + dx
        o += cout * cin * 9
        out.append(w[o:o + cout])
        o += cout
    return out


def _pack(tensors):
    parts = []
    for k in range(0, 6, 2):
        W, b = tensors[k], tensors[k + 1]
        cout, cin = W.shape[:2]
        parts.append(W.reshape(cout, cin, 9).permute(0, 2, 1).reshape(-1))
        parts.append(b)
    return torch.cat(parts).numpy()


def _bf(t):
    return t + (t.to(torch.bfloat16).to(torch.float32) - t).detach()


def _reference(x, w, y=None, pos_weight=1.0):
    params = [p.clone().requires_grad_(True) for p in _unpack(w)]
    W1, b1, W2, b2, W3, b3 = params
    xb = _bf(torch.tensor(x))
    a1 = _bf(F.relu(F.conv2d(xb, _bf(W1), b1)))
    a2 = _bf(F.relu(F.conv2d(a1, _bf(W2), b2, padding=1)))
    z3 = F.conv2d(a2, _bf(W3), b3, padding=1)
    logits = z3.flatten(2).max(dim=2).values
    if y is None:
        return logits.detach().numpy(), None, None
    n = x.shape[0]
    loss = F.binary_cross_entropy_with_logits(logits, torch.tensor(y), reduction="sum",
                                              pos_weight=torch.tensor([pos_weight, pos_weight])) / n
    loss.backward()
    return logits.detach().numpy(), _pack([p.grad for p in params]), float(loss)


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 3, 18, 18), dtype=np.float32)
    yy, xx = np.mgrid[0:18, 0:18]
    for k in range(n):   # tanh interfaces at random angles and offsets, random velocities
        a, c = rng.uniform(0, 2 * np.pi), rng.uniform(-12, 12)
        x[k, 0] = 0.5 + 0.5 * np.tanh(((xx - 8.5) * np.cos(a) + (yy - 8.5) * np.sin(a) - c) / 2.5)
        x[k, 1, 1:17, 1:17] = rng.normal(0, 0.3, (16, 16))
        x[k, 2, 1:17, 1:17] = rng.normal(0, 0.3, (16, 16))
    y = (rng.random((n, 2)) < 0.3).astype(np.float32)
    return x, y


def test_forward_matches_torch(amrlbm):
    w = amrlbm.cnn_init(3)
    x, _ = _inputs(32, 5)
    got = amrlbm.cnn_forward(x, w)
    ref, _, _ = _reference(x, w)
    err = np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-6)
    print(f"[cnn] forward: max |logit| {np.max(np.abs(ref)):.3e}, max rel err {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("pos_weight", [1.0, 5.0])
def test_gradient_matches_torch(amrlbm, pos_weight):
    w = amrlbm.cnn_init(7)
    x, y = _inputs(24, 11)
    g, loss = amrlbm.cnn_gradient(x, y, w, pos_weight)
    _, gref, lref = _reference(x, w, y, pos_weight)
    scale = max(float(np.max(np.abs(gref))), 1e-12)
    err = np.max(np.abs(g - gref)) / scale
    print(f"[cnn] gradient (pos_weight {pos_weight}): loss {loss:.6f} vs {lref:.6f}, "
          f"max |g| {scale:.3e}, max err / max |g| {err:.2e}")
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    assert err <= 1e-4


def test_training_learns_an_interface_rule(amrlbm):
    """Target: channel 0 set iff an interface (0.1 < C < 0.9) crosses the
    block's left half, channel 1 iff it crosses the right half. Training on
    the GPU must bring the loss down and classify held-out blocks."""
    x, _ = _inputs(3000, 21)
    inter = (x[:, 0] > 0.1) & (x[:, 0] < 0.9)
    y = np.stack([inter[:, 1:17, 1:9].any(axis=(1, 2)), inter[:, 1:17, 9:17].any(axis=(1, 2))], 1).astype(np.float32)
    w0 = amrlbm.cnn_init(1)
    w, hist = amrlbm.cnn_train(x[:2500], y[:2500], w0, epochs=30, batch=128, lr=3e-3)
    acc = np.mean((amrlbm.cnn_forward(x[2500:], w) > 0) == (y[2500:] > 0.5))
    print(f"[cnn] training: loss {hist[0]:.4f} -> {hist[-1]:.4f}, held-out accuracy {acc:.3f}")
    assert hist[-1] < 0.5 * hist[0]
    assert acc >= 0.9
