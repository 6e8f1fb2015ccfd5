"""The FP64 CNN reference (oracle/cnn_ref.py) pinned by central finite
differences: its analytic backward must match numerical derivatives of its
own forward loss (no bf16 rounding, float64)."""
import numpy as np

import cnn_ref as R


def test_reference_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    n = sum(9 * ci * co + co for ci, co in zip(R.CIN, R.COUT))
    params = rng.normal(0, 0.15, n)
    # zero the padded channels like the product does (input 5 real, output 3 real)
    o = 0
    for li, (ci, co) in enumerate(zip(R.CIN, R.COUT)):
        w = params[o:o + 9 * co * ci].reshape(9, co, ci)
        if li == 0:
            w[:, :, 5:] = 0
        if li == 3:
            w[:, 3:, :] = 0
        o += 9 * co * ci
        if li == 3:
            params[o + 3:o + co] = 0
        o += co
    x = np.zeros((1, 8, 8, 16))
    x[..., :5] = rng.uniform(0, 1, (1, 8, 8, 5))
    target = rng.uniform(0, 1, (1, 8, 8, 3))
    mask = (rng.random((1, 8, 8)) < 0.5).astype(np.float64)
    loss, g = R.loss_and_grads(params, x, target, mask, round_bf16=False)
    idx = rng.choice(np.nonzero(g)[0], 60, replace=False)
    eps = 1e-6
    for i in idx:
        p1 = params.copy(); p1[i] += eps
        p2 = params.copy(); p2[i] -= eps
        l1, _ = R.loss_and_grads(p1, x, target, mask, round_bf16=False)
        l2, _ = R.loss_and_grads(p2, x, target, mask, round_bf16=False)
        num = (l1 - l2) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])


def test_reference_forward_matches_torch_conv2d():
    """The forward's tap orientation, zero padding, bias and ReLU against
    torch.nn.functional.conv2d (cross-correlation, padding 1) in float64:
    tap t = 3 dy + dx of the [9][cout][cin] layout is weight[co, ci, dy, dx]."""
    import torch
    import torch.nn.functional as F
    rng = np.random.default_rng(1)
    n = sum(9 * ci * co + co for ci, co in zip(R.CIN, R.COUT))
    params = rng.normal(0, 0.2, n)
    x = rng.uniform(-1, 1, (2, 11, 13, 16))
    out, _ = R.forward(params, x, round_bf16=False)
    h = torch.from_numpy(x).permute(0, 3, 1, 2)
    for i, (w, b) in enumerate(R.split(params)):
        wt = torch.from_numpy(np.ascontiguousarray(w.reshape(3, 3, w.shape[1], w.shape[2]).transpose(2, 3, 0, 1)))
        h = F.conv2d(h, wt, torch.from_numpy(b), padding=1)
        if i < 3:
            h = torch.relu(h)
    ref = h.permute(0, 2, 3, 1).numpy()
    assert np.abs(out - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    # a transposed tap (dy <-> dx) breaks the agreement
    bad = params.copy()
    w0 = bad[:9 * 32 * 16].reshape(3, 3, 32, 16)
    bad[:9 * 32 * 16] = w0.transpose(1, 0, 2, 3).reshape(-1)
    out_bad, _ = R.forward(bad, x, round_bf16=False)
    assert np.abs(out_bad - ref).max() > 1e-3
