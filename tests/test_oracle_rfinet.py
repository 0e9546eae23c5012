"""The RFI network's CPU restatement pinned independently: its gradient
against central finite differences of its own loss (float64), BF16 rounding
against hand-worked values, and the features against closed forms."""
import numpy as np

from oracle import rfinet as R


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = rng.standard_normal(R.nparam()) * 0.2
    feat = rng.standard_normal((2, R.F, 23))
    labels = (rng.random((2, 23)) < 0.3).astype(np.uint8)
    loss, g = R.loss_grad(theta, feat, labels, pos_weight=3.0)
    idx = rng.choice(R.nparam(), 60, replace=False)
    # include biases and every layer
    offs = np.cumsum([0] + [co * ci * R.K + co for co, ci in R.SHAPES])
    idx = np.concatenate([idx, offs[1:] - 1, offs[:-1]])
    h = 1e-6
    for i in idx:
        tp = theta.copy()
        tp[i] += h
        tm = theta.copy()
        tm[i] -= h
        fd = (R.loss_grad(tp, feat, labels, 3.0)[0] - R.loss_grad(tm, feat, labels, 3.0)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7 + 2 ** -8, -2.5, 3.0e-3], np.float32)
    r = R.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # tie, even mantissa stays
    assert r[2] == np.float32(1 + 2 ** -7)  # above the tie rounds up
    assert r[3] == np.float32(1 + 2 ** -6)  # tie with odd mantissa rounds up to even
    assert r[4] == -2.5
    assert (r.view(np.uint32) & 0xFFFF).max() == 0


def test_features_closed_forms():
    """Constant power blocks: SK - 1 = -1 exactly (S2 = S1^2/M); the log
    power feature is zero-mean per channel; the difference feature is the
    block-to-block change."""
    M, nblk, nchan = 64, 3, 2
    p = np.array([[1.0, 4.0], [2.0, 4.0], [4.0, 4.0]], np.float32)  # power per block, channel
    sums = np.zeros((nblk, nchan, 1, 2), np.float32)
    sums[:, :, 0, 0] = M * p
    sums[:, :, 0, 1] = M * p * p
    counts = np.full((nblk, nchan), M, np.uint32)
    f = R.features(sums, counts)
    assert np.allclose(f[:, 1], np.float32(M + 1) / (M - 1) * 0 - 1, atol=2e-6)
    assert np.allclose(f[:, 0].mean(axis=0), 0, atol=1e-6)
    assert np.allclose(f[1:, 3, 0], np.log(2), atol=1e-6) and np.allclose(f[:, 3, 1], 0)


def test_forward_and_gradient_match_torch():
    """The restated forward (four 'same' zero-padded 1-D convolutions with
    ReLU between them) against torch.nn.functional.conv1d, and the
    restated loss and gradient against torch's autograd of
    binary_cross_entropy_with_logits with the positive-class weight, all in
    float64 on the CPU."""
    import torch
    import torch.nn.functional as Fn
    rng = np.random.default_rng(4)
    theta = rng.standard_normal(R.nparam()) * 0.2
    feat = rng.standard_normal((3, R.F, 37))
    labels = (rng.random((3, 37)) < 0.25).astype(np.uint8)
    z, _ = R.forward(theta, feat)
    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    params, o = [], 0
    for co, ci in R.SHAPES:
        w = t[o:o + co * ci * R.K].view(co, ci, R.K)
        o += co * ci * R.K
        params.append((w, t[o:o + co]))
        o += co
    a = torch.tensor(feat, dtype=torch.float64)
    for l, (w, b) in enumerate(params):
        a = Fn.conv1d(a, w, b, padding=R.P)
        if l < R.L - 1:
            a = torch.relu(a)
    zt = a[:, 0]
    assert np.allclose(z, zt.detach().numpy(), rtol=1e-12, atol=1e-12)
    y = torch.tensor(labels, dtype=torch.float64)
    loss_t = Fn.binary_cross_entropy_with_logits(zt, y, pos_weight=torch.tensor(3.0, dtype=torch.float64))
    loss_t.backward()
    loss, g = R.loss_grad(theta, feat, labels, pos_weight=3.0)
    assert abs(loss - loss_t.item()) < 1e-12
    assert np.allclose(g, t.grad.numpy(), rtol=1e-10, atol=1e-14)
