"""Autotuner for the ray kernel's launch knobs.

Knobs: ray hand-out order (svo_set_ray_spread: the size of the ray groups
spread across waves, 64 = list order; lanes per wave, svo_set_ray_lanes, is
applied when a setting names it but is not searched by default: fewer lanes
helped short windows and hurt long ones) and, through svo_set_tuning, the work-queue refill threshold (idle
lanes per refill), persistent workgroups per CU, traversal steps between
refill checks, the LDS node-cache depth (octree words of the breadth-first
top levels kept in LDS) and the shape of the blocks in which rectangle rays
are handed out (svo_set_ray_block); full frames also search the trip count
that puts a screen block on the long-rays-first list (svo_set_long_first),
and learned frames the CNN's workgroups alongside the ray launch
(svo_set_cnn_overlap); warp and learned frames also search the pretrace's
result sets, its mode (long blocks or predicted holes) and block threshold
(svo_set_pretrace).  Every candidate replays the same
frame window from the same starting state (the window starts with the two
clearing frames), so candidates see identical work; the frames are issued
back to back and the score is the mean time between the completions of
successive timed frames (timing mode 3: the frame period the GPU sustains,
which frames/s follows).  Results do not depend on the knobs (the tests
check bit-exactness under several settings).
"""
from __future__ import annotations

import itertools
import statistics
import time

DEFAULT_SPACE = {
    "spread": [1, 64],
    "refill": [8, 16, 32, 48],
    "wg_per_cu": [2, 4, 3, 8, 16],  # the second entry starts the descent
    "iters": [8, 16, 64],
    "lds_cache_words": [0, 96, 736, 2048, 4096],
    "ray_block": [16, 8, 32, 64],  # the second entry starts the descent: 8x8
}


def _frame_fn(ctx, mode, cnn, max_cnn_holes):
    if mode == "warp":
        return ctx.frame_warp
    if mode == "full":
        return ctx.frame_full
    if mode == "learned":
        return lambda p: ctx.frame_warp_learned(cnn, p, max_cnn_holes)
    raise ValueError(f"unknown frame mode {mode!r}")


def _frame_time(ctx, params, warmup, mode, cnn=None, max_cnn_holes=64):
    frame = _frame_fn(ctx, mode, cnn, max_cnn_holes)
    ctx.enable_timing(False)
    for p in params[:warmup]:
        frame(p)
    ctx.sync()
    ctx.enable_timing(mode=3)  # completion events, frames back to back
    for p in params[warmup:]:
        frame(p)
    ts = ctx.frame_intervals()
    ctx.enable_timing(False)
    return float(statistics.mean(ts))  # frames/s is set by the mean


def _search(space, coordinate, measure):
    keys = list(space)
    best = {k: space[k][min(1, len(space[k]) - 1)] for k in keys}
    if coordinate:
        best_ms = measure(best)
        for _ in range(2):
            for k in keys:
                for v in space[k]:
                    if v == best[k]:
                        continue
                    cfg = dict(best, **{k: v})
                    ms = measure(cfg)
                    if ms < best_ms:
                        best, best_ms = cfg, ms
    else:
        best_ms = float("inf")
        for vals in itertools.product(*[space[k] for k in keys]):
            cfg = dict(zip(keys, vals))
            ms = measure(cfg)
            if ms < best_ms:
                best, best_ms = cfg, ms
    return best, best_ms


# warp frames also search the pretrace: its sets, its mode (pre_pred: 0 =
# long blocks, k = holes predicted from frame g-k), block threshold, the
# hardware queues it runs on (pre_qskip: every fourth queue of its pool left
# out), its launch's workgroups per CU (pre_wg) and the waves of their own
# that its longest rays get (pre_split: known trips at least this many,
# pre_split_lanes rays per wave)
WARP_SPACE = dict(DEFAULT_SPACE, pre_sets=[4, 3, 5, 0], pre_pred=[0, 2, 3], pre_long=[256, 1, 128, 384],
                  pre_qskip=[1, 3, 0, 2, 4], pre_wg=[2, 1],
                  pre_split=[1024, 2048, 0], pre_split_lanes=[4, 1])
# full frames also search the long-ray-first threshold (svo_set_long_first)
FULL_SPACE = dict(DEFAULT_SPACE, long_trips=[0, 1024, 512, 2048])
# learned frames also search the CNN's workgroups alongside the ray launch
LEARNED_SPACE = dict(DEFAULT_SPACE, cnn_wgs=[256, 128, 96, 160, 192, 512], pre_sets=[4, 3, 5, 0],
                     pre_pred=[0, 2, 3], pre_long=[256, 1, 128, 384], pre_qskip=[1, 3, 0, 2, 4], pre_wg=[2, 1],
                     pre_split=[1024, 2048, 0], pre_split_lanes=[4, 1])


def apply(ctx, cfg):
    """Apply a knob setting (as returned by tune / tune_batch) to a context."""
    # knobs missing from cfg keep their current settings (-1)
    ctx.set_tuning(cfg.get("refill", -1), cfg.get("wg_per_cu", -1), cfg.get("iters", -1), cfg.get("lds_cache_words", -1))
    if "spread" in cfg:
        ctx.set_ray_spread(int(cfg["spread"]), True)
    if "lanes" in cfg:
        ctx.set_ray_lanes(int(cfg["lanes"]))
    if "long_trips" in cfg:
        ctx.set_long_first(int(cfg["long_trips"]))
    if "ray_block" in cfg:
        ctx.set_ray_block(int(cfg["ray_block"]))
    if "cnn_wgs" in cfg:
        ctx.set_cnn_overlap(int(cfg["cnn_wgs"]))
    if any(k in cfg for k in ("pre_sets", "pre_long", "pre_pred", "pre_qskip", "pre_wg", "pre_split",
                              "pre_split_lanes")):
        ctx.set_pretrace(sets=cfg.get("pre_sets"), long_trips=cfg.get("pre_long"), predict=cfg.get("pre_pred"),
                         queue_skip=cfg.get("pre_qskip"), wg_per_cu=cfg.get("pre_wg"),
                         split_trips=cfg.get("pre_split"), split_lanes=cfg.get("pre_split_lanes"))


def tune(ctx, params, warmup=4, mode="warp", space=None, coordinate=True, log=None, cnn=None, max_cnn_holes=64):
    """Pick the fastest knob setting for this context's scene/resolution.

    params: FrameParams list (params[0] must have frame 0).  mode: "warp",
    "full" or "learned" (svo_frame_warp_learned with `cnn`).  With
    coordinate=True the knobs are tuned one at a time (coordinate descent,
    two sweeps); otherwise the full grid is measured.  Applies and returns
    (best_setting, best_ms, trials)."""
    if space is None:
        space = {"full": FULL_SPACE, "learned": LEARNED_SPACE, "warp": WARP_SPACE}.get(mode, DEFAULT_SPACE)
    space = dict(space)
    trials = []

    def measure(cfg):
        apply(ctx, cfg)
        ms = _frame_time(ctx, params, warmup, mode, cnn, max_cnn_holes)
        trials.append((dict(cfg), ms))
        if log:
            log(f"[autotune] {cfg} -> {ms:.4f} ms")
        return ms

    best, best_ms = _search(space, coordinate, measure)
    apply(ctx, best)
    return best, best_ms, trials


def tune_batch(batch, frames, warmup=4, space=None, coordinate=True, log=None):
    """Tune the shared ray launch of a Batch (its first view's knobs).

    frames: list of per-frame FrameParams lists, one entry per view, starting
    at frame 0.  The score is the mean wall time of one batched frame
    (launch + device completion); returns (best_setting, best_ms, trials)."""
    space = dict(DEFAULT_SPACE if space is None else space)
    ctx0 = batch.contexts[0]
    trials = []

    def measure(cfg):
        apply(ctx0, cfg)
        for ps in frames[:warmup]:
            batch.frame_warp(ps)
        batch.sync()
        ts = []
        for ps in frames[warmup:]:
            t0 = time.perf_counter()
            batch.frame_warp(ps)
            batch.sync()
            ts.append((time.perf_counter() - t0) * 1e3)
        ms = statistics.mean(ts)
        trials.append((dict(cfg), ms))
        if log:
            log(f"[autotune batch] {cfg} -> {ms:.4f} ms")
        return ms

    best, best_ms = _search(space, coordinate, measure)
    apply(ctx0, best)
    return best, best_ms, trials
