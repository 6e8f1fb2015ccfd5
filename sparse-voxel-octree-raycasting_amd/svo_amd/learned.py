"""Training and evaluation of the learned hole filler on frame pairs the
pipeline generates itself: a warp context renders the gated warp frame
without the CNN (its CNN input is plane 0 as the gate leaves it, and the
holes left after its rays are the pixels the CNN fills), a second context
renders the full raycast of the same camera (the target image), and
the CNN takes one Adam step on random crops of each pair.
"""
from __future__ import annotations

import os

import numpy as np

from . import Cnn, Context, image_rgb, psnr
from . import workloads as W

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "holefill.npz")


def _contexts(name):
    words, root, depth, _ = W.octree(name)
    rx, ry = W.WORKLOADS[name]["res"]
    ctxs = []
    for _ in range(2):
        c = Context(0)
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
        ctxs.append(c)
    return ctxs


def train(cnn, plan, n_crops=16, crop=64, lr=1e-3, lr_end=1e-4, log=None, log_every=200, max_cnn_holes=64):
    """plan: list of (workload name, path seed offset, n_frames).  Frames are
    replayed along each seeded path; one step per frame pair.  Returns the
    per-step losses (every log_every steps, others None)."""
    losses = []
    total = sum(n for _, _, n in plan)
    step = 0
    for name, seed_off, n_frames in plan:
        warp, full = _contexts(name)
        w = W.WORKLOADS[name]
        rows = W.path(name, w["frames"]) if seed_off == 0 else _path(name, seed_off)
        for f in range(n_frames):
            r = f % len(rows)
            p = W.frame_params(name, rows, r, frame_index=f)
            warp.frame_warp_learned(None, p, max_cnn_holes)
            full.frame_full(p)
            a = step / max(1, total - 1)
            cur = lr * (lr_end / lr) ** a
            want = (step % log_every == 0) or step == total - 1
            loss = cnn.train_step(warp, full, p, n_crops, crop, seed=step, lr=cur, want_loss=want)
            losses.append(loss)
            if want and log:
                log(f"[train] step {step} {name} frame {f} lr {cur:.2e} loss {loss:.5f}")
            step += 1
        warp.sync()
        full.sync()
        warp.close()
        full.close()
    return losses


def _path(name, seed_off):
    from . import camera_path
    w = W.WORKLOADS[name]
    return camera_path(w["path_seed"] + seed_off, w["depth"], w["frames"], w["vox"], w["yaw"], w["height"],
                       w["pitch"])


def evaluate(cnn, name, seed_off=1000, n_frames=40, skip=8, max_cnn_holes=64, gate_depth=0.0, detail=False):
    """PSNR (dB, vs the full raycast) of the plain warp (fillhole2), the gate
    alone, and the gate + CNN, over n_frames of a held-out path; mean over the
    frames after the first `skip` that differ from the full raycast, on all
    pixels and on the hole pixels (NaN when every frame is identical).
    gate_depth: the gate's depth rule (svo_set_gate_depth) for both gated
    modes.  detail=True adds a third dict per mode: the PSNR of the mean
    squared error over all those frames ("seq_db", None when it is zero) and
    the number of frames identical to the full raycast."""
    words, root, depth, _ = W.octree(name)
    rx, ry = W.WORKLOADS[name]["res"]
    rows = _path(name, seed_off)
    ctx = {k: Context(0) for k in ("plain", "gate", "learned", "full")}
    for c in ctx.values():
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
    ctx["gate"].set_gate_depth(gate_depth)
    ctx["learned"].set_gate_depth(gate_depth)
    modes = ("plain", "gate", "learned")
    res = {k: [] for k in modes}
    hole_res = {k: [] for k in modes}
    sq = {k: 0.0 for k in modes}
    same = {k: 0 for k in modes}
    n_eval = 0
    for f in range(n_frames):
        p = W.frame_params(name, rows, f % len(rows), frame_index=f)
        ctx["plain"].frame_warp(p)
        ctx["gate"].frame_warp_learned(None, p, max_cnn_holes)
        ctx["learned"].frame_warp_learned(cnn, p, max_cnn_holes)
        ctx["full"].frame_full(p)
        if f < skip:
            continue
        n_eval += 1
        ref = image_rgb(ctx["full"].image())
        holes = ctx["learned"].screen(2) == 0xFFFFFF00
        for k in res:
            img = image_rgb(ctx[k].image())
            d = img.astype(np.float64) - ref.astype(np.float64)
            sq[k] += float(np.mean(d * d))
            same[k] += int(not d.any())
            res[k].append(psnr(img, ref))
            if holes.any():
                hole_res[k].append(psnr(img[holes], ref[holes]))
    for c in ctx.values():
        c.close()

    def fin(v):
        v = [x for x in v if np.isfinite(x)]
        return float(np.mean(v)) if v else float("nan")

    out = ({k: fin(v) for k, v in res.items()}, {k: fin(v) for k, v in hole_res.items()})
    if not detail:
        return out
    seq = {k: (10.0 * np.log10(255.0 ** 2 / (sq[k] / max(1, n_eval))) if sq[k] > 0 else None) for k in modes}
    return out + ({k: {"seq_db": None if seq[k] is None else float(seq[k]), "identical_frames": same[k],
                       "frames": n_eval} for k in modes},)


def save(cnn, path=WEIGHTS, **meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, params=cnn.params(), **{k: np.asarray(v) for k, v in meta.items()})


def load(cnn, path=WEIGHTS):
    z = np.load(path, allow_pickle=False)
    cnn.set_params(z["params"])
    return {k: z[k] for k in z.files if k != "params"}
