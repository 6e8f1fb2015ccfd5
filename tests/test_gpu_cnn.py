"""GPU BF16 MFMA CNN kernels vs the FP64 CPU reference (oracle/cnn_ref.py,
itself pinned by finite differences): forward outputs and every parameter
gradient.  Tolerances are stated per check (bf16 activations and gradients,
FP32 accumulation)."""
import numpy as np
import pytest

import cnn_ref as R
import svo_amd as S

pytestmark = pytest.mark.gpu


def _batch(rng, B, H, W):
    x = np.zeros((B, H, W, 16), np.float32)
    x[..., :3] = rng.uniform(0, 1, (B, H, W, 3))
    x[..., 3] = rng.uniform(0, 1, (B, H, W))
    x[..., 4] = rng.random((B, H, W)) < 0.3
    x[..., :3] *= 1 - x[..., 4:5]
    bits = R.bf16_bits(x)
    return bits, R.bits_to_f64(bits)


@pytest.fixture(scope="module")
def cnn():
    c = S.Cnn(0, seed=7)
    yield c
    c.close()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_reference(cnn):
    rng = np.random.default_rng(1)
    bits, x = _batch(rng, 3, 24, 16)
    p = cnn.params()
    out = cnn.forward(bits)
    ref, _ = R.forward(p.astype(np.float64), x)
    # bf16 inter-layer activations are rounded identically; the residual is
    # FP32 accumulation order (and rare bf16 rounding-boundary flips)
    assert rel(out[..., :3], ref[..., :3]) < 5e-3
    assert np.all(out[..., 3:] == 0.0)


def test_gradients_match_reference(cnn):
    rng = np.random.default_rng(2)
    bits, x = _batch(rng, 2, 16, 24)
    target = rng.uniform(0, 1, (2, 16, 24, 3)).astype(np.float32)
    mask = (rng.random((2, 16, 24)) < 0.4).astype(np.float32)
    p = cnn.params()
    g, loss = cnn.gradients(bits, target, mask)
    lref, gref = R.loss_and_grads(p.astype(np.float64), x, target.astype(np.float64), mask.astype(np.float64))
    assert abs(loss - lref) <= 1e-3 * lref
    o = 0
    for li, (ci, co) in enumerate(zip(R.CIN, R.COUT)):
        nw = 9 * ci * co
        gw, gwr = g[o:o + nw], gref[o:o + nw]
        gb, gbr = g[o + nw:o + nw + co], gref[o + nw:o + nw + co]
        o += nw + co
        # backward keeps dZ in bf16 (8-bit mantissa): 3% relative L2 per layer
        assert rel(gw, gwr) < 3e-2, (li, rel(gw, gwr))
        assert rel(gb, gbr) < 3e-2, (li, rel(gb, gbr))
    # padded channels receive no gradient
    assert np.all(g[:9 * 32 * 16].reshape(9, 32, 16)[:, :, 5:] == 0)


def test_gradients_are_deterministic(cnn):
    """The weight and bias gradients are summed from per-crop and per-block
    partials in a fixed order: the same batch gives bit-identical gradients
    every time (16 crops of 64x64, the training shape)."""
    rng = np.random.default_rng(3)
    bits, _ = _batch(rng, 16, 64, 64)
    target = rng.uniform(0, 1, (16, 64, 64, 3)).astype(np.float32)
    mask = (rng.random((16, 64, 64)) < 0.4).astype(np.float32)
    g0, _ = cnn.gradients(bits, target, mask)
    for _ in range(3):
        g1, _ = cnn.gradients(bits, target, mask)
        assert np.array_equal(g0.view(np.uint32), g1.view(np.uint32))
    assert np.abs(g0).max() > 0


@pytest.mark.parametrize("mode", [1, 2])
def test_fused_tile_forward_matches_reference(mode):
    """The fused per-tile inference kernel (all four layers in LDS, layer 1 on
    16x16x16 MFMAs, layers 2-4 on 16x16x32 MFMAs) against the FP64 network on
    full 24x24 crops, compared on the centre 16x16 it returns."""
    rng = np.random.default_rng(5)
    c = S.Cnn(0, seed=11)
    try:
        p = c.params()
        o = 0
        for ci, co in zip(R.CIN, R.COUT):  # non-zero biases on the real output channels
            o += 9 * ci * co
            p[o:o + co] = rng.uniform(-0.2, 0.2, co) * (np.arange(co) < (3 if co == 16 else co))
            o += co
        c.set_params(p)
        c.set_fused(mode=mode)
        bits, x = _batch(rng, 5, 24, 24)
        out = c.forward_tiles(bits)
        ref, _ = R.forward(c.params().astype(np.float64), x)
        assert rel(out, ref[:, 4:20, 4:20, :3]) < 5e-3
    finally:
        c.close()
