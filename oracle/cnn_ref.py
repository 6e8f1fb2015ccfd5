"""FP64 CPU reference of the learned hole filler (test infrastructure): the
4-layer 3x3 CNN of the product's svo_cnn.hip, forward and backward, with the
same bf16 rounding of weights and inter-layer activations.  Its backward is
itself pinned by central finite differences (tests/test_cnn_reference.py).

Parameter vector layout (shared with svo_cnn_get_params): per layer the
weights [9 taps][cout][cin] then the bias [cout]; channels 16/32/32/32 in,
32/32/32/16 out.
"""
import numpy as np

CIN = (16, 32, 32, 32)
COUT = (32, 32, 32, 16)


def bf16_round(a):
    """float -> nearest-even bfloat16, returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bits_to_f64(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def split(params):
    out, o = [], 0
    for ci, co in zip(CIN, COUT):
        w = params[o:o + 9 * co * ci].reshape(9, co, ci)
        o += 9 * co * ci
        b = params[o:o + co]
        o += co
        out.append((np.asarray(w, np.float64), np.asarray(b, np.float64)))
    return out


def conv_same(x, w):
    """x [B,H,W,ci], w [9,co,ci] -> [B,H,W,co], zero padding."""
    B, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((B, H, W, w.shape[1]))
    for t in range(9):
        dy, dx = divmod(t, 3)
        out += np.einsum("bhwc,oc->bhwo", xp[:, dy:dy + H, dx:dx + W, :], w[t])
    return out


def conv_same_wgrad(x, dz):
    """dW[t,co,ci] = sum dz[p,co] x[p + tap - 1, ci]."""
    B, H, W, ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    g = np.zeros((9, dz.shape[3], ci))
    for t in range(9):
        dy, dx = divmod(t, 3)
        g[t] = np.einsum("bhwo,bhwc->oc", dz, xp[:, dy:dy + H, dx:dx + W, :])
    return g


def conv_same_dgrad(dz, w):
    """dX[p,ci] = sum_t sum_co w[t,co,ci] dz[p - tap + 1, co]."""
    B, H, W, _ = dz.shape
    dzp = np.pad(dz, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((B, H, W, w.shape[2]))
    for t in range(9):
        dy, dx = divmod(t, 3)
        out += np.einsum("bhwo,oc->bhwc", dzp[:, 2 - dy:2 - dy + H, 2 - dx:2 - dx + W, :], w[t])
    return out


def forward(params, x, round_bf16=True):
    """x: float [B,H,W,16] (already bf16-representable).  Returns (out, acts)."""
    layers = split(params)
    acts = [x]
    h = x
    for i, (w, b) in enumerate(layers):
        wq = bf16_round(w) if round_bf16 else w
        z = conv_same(h, wq) + b
        if i < 3:
            h = np.maximum(z, 0.0)
            if round_bf16:
                h = bf16_round(h)
            acts.append(h)
        else:
            h = z
    return h, acts


def loss_and_grads(params, x, target, mask, round_bf16=True):
    layers = split(params)
    out, acts = forward(params, x, round_bf16)
    cnt = mask.sum()
    d = out[..., :3] - target
    loss = float((mask[..., None] * d * d).sum() / (3 * cnt))
    dz = np.zeros_like(out)
    dz[..., :3] = 2 * mask[..., None] * d / (3 * cnt)
    grads = [None] * 4
    for i in range(3, -1, -1):
        w, b = layers[i]
        wq = bf16_round(w) if round_bf16 else w
        grads[i] = (conv_same_wgrad(acts[i], dz), dz.sum(axis=(0, 1, 2)))
        if i > 0:
            dz = conv_same_dgrad(dz, wq) * (acts[i] > 0)
    flat = np.concatenate([np.concatenate([g.reshape(-1), gb]) for g, gb in grads])
    return loss, flat
