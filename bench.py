#!/usr/bin/env python
"""Headline benchmark: frames/s of the warp + fill pipeline at 1920x1024 on the
depth-11 flythrough (vs full per-frame raycast), p50/p99 frame ms, Mrays/s,
roofline of the dominant kernel and the CPU-oracle baseline.

    python bench.py [--gpus 1] [--steps 100] [--warmup 10] [--no-cpu]

A step is one frame of the pipeline over the 1920x1024 screen with the scene
resident in HBM.  Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

# The warp frames' pretrace runs each of its result sets on a stream of its
# own; HIP maps a process's streams onto at most GPU_MAX_HW_QUEUES hardware
# queues (4 by default) and streams that share a queue run one after the
# other.  Read at the first HIP call, so set before the library loads.
HW_QUEUES = 16
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < HW_QUEUES:
    os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sparse-voxel-octree-raycasting_amd"))

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
# fused CNN tile network, per 16x16 tile: useful FLOPs = 2 * (484*9*5*32 +
# 400*9*32*32 + 324*9*32*32 + 256*9*32*3) (real channels on the 22/20/18/16
# pixel-square regions the centre needs); issued = the MFMAs the kernel runs
# (32/26/22/16 M-tiles, channels padded to 16/32)
CNN_TILE_FLOPS = 2 * (484 * 9 * 5 * 32 + 400 * 9 * 32 * 32 + 324 * 9 * 32 * 32 + 256 * 9 * 32 * 3)
CNN_TILE_MFMA_FLOPS = 32 * 2 * 9 * 8192 + 26 * 2 * 9 * 16384 + 22 * 2 * 9 * 16384 + 16 * 9 * 16384
METRIC = "frames/s & p99 frame ms at 1920x1024 (warp+fill vs full raycast); Mrays/s"
TUNE_SEED_OFF = 500  # the autotuner's held-out camera paths
CPU_SHARE = 16  # host CPUs per GPU on the measurement boxes


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else float("nan")


def num(v, nd=4):
    """round for the JSON line; None for NaN"""
    return None if v != v else round(v, nd)


def run_gpu(args, name):
    import svo_amd as S
    from svo_amd import autotune as A
    from svo_amd import workloads as W

    w = W.WORKLOADS[name]
    t = time.time()
    xyz, rgba = W.voxels(name)
    t_scene = time.time() - t
    t = time.perf_counter()
    words, root, depth, nvox = W.octree(name)
    t_host = time.perf_counter() - t
    log(f"[bench] scene {name}: {nvox} voxels ({t_scene:.1f}s) -> {len(words)} words in {t_host:.2f}s (host builder)")
    # the same build on the GPU (identical words; checked here as well as in the tests)
    S.build_octree_gpu(xyz[:1000], rgba[:1000], w["depth"])  # warm-up: library code load
    t = time.perf_counter()
    gwords, groot = S.build_octree_gpu(xyz, rgba, w["depth"])
    t_gpu = time.perf_counter() - t
    build = dict(voxels=int(nvox), words=int(len(words)), host_s=round(t_host, 3), gpu_s=round(t_gpu, 3),
                 identical=bool(groot == root and np.array_equal(gwords, words)))
    del gwords
    log(f"[bench] GPU builder: {build}")
    rx, ry = w["res"]
    n_frames = w["frames"]
    total = args.warmup + args.steps
    if total > n_frames:
        raise SystemExit(f"warmup + steps must be <= {n_frames} (path length)")
    rows = W.path(name, n_frames)
    params = [W.frame_params(name, rows, f) for f in range(total)]

    warp = S.Context(0)
    full = S.Context(0)
    for c in (warp, full):
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
    res = {"octree_build": build}
    # the autotuner replays frames of a held-out path (same scene and motion
    # bounds, another seed), so the knobs are not picked on the timed frames
    trows = W.path(name, 70, seed_off=TUNE_SEED_OFF)
    tp = [W.frame_params(name, trows, f) for f in range(70)]
    if args.tune:
        from svo_amd import autotune as A
        for c, mode in ((warp, "warp"), (full, "full")):
            best, ms, _ = A.tune(c, tp, warmup=10, mode=mode)
            res[f"tuned_{mode}"] = dict(best, ms=round(ms, 4))
            log(f"[bench] autotune {mode}: {best} -> {ms:.4f} ms/frame")
    if not args.pretrace:
        warp.set_pretrace(sets=0)

    # ---- pass A: throughput, frames back to back (no per-frame host sync);
    # an event at each frame's end gives the frame-to-frame completion times
    for f in range(args.warmup):
        warp.frame_warp(params[f])
    warp.sync()
    warp.ray_stats(reset=True)
    warp.pretrace_stats(reset=True)
    warp.enable_timing(mode=3)
    t0 = time.perf_counter()
    for f in range(args.warmup, total):
        warp.frame_warp(params[f])
    warp.sync()
    t_warp = time.perf_counter() - t0
    res["warp_intervals"] = warp.frame_intervals()
    warp.enable_timing(False)
    rays_traced, longest_frame = warp.ray_stats(reset=True)
    pre_rays, pre_merged, pre_longest = warp.pretrace_stats(reset=True)
    rays_warp = rays_traced + pre_merged  # the frames' hole and tile rays
    cfg = warp.get_pretrace()
    res["pretrace"] = {"sets": cfg.sets, "predict": cfg.predict, "long_trips": cfg.long_trips, "wg_per_cu": cfg.wg_per_cu,
                       "queue_skip": cfg.queue_skip,
                       "only_queued": cfg.only_queued, "hip_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                       "rays_traced_ahead_per_frame": round(pre_rays / args.steps),
                       "frame_rays_taken_per_frame": round(pre_merged / args.steps),
                       "frame_rays_traced_per_frame": round(rays_traced / args.steps),
                       "longest_ray_traced_ahead": pre_longest, "longest_ray_traced_in_frame": longest_frame}
    log(f"[bench] warp: {args.steps} frames in {t_warp * 1e3:.1f} ms; pretrace {res['pretrace']}")

    t_full, rays_full = float("nan"), 0
    if args.legs == "all":
        for f in range(min(args.warmup, 3)):
            full.frame_full(params[f])
        full.sync()
        full.ray_count(reset=True)
        full.enable_timing(mode=3)
        t0 = time.perf_counter()
        for f in range(args.warmup, total):
            full.frame_full(params[f])
        full.sync()
        t_full = time.perf_counter() - t0
        res["full_intervals"] = full.frame_intervals()
        full.enable_timing(False)
        rays_full = full.ray_count(reset=True)
        log(f"[bench] full: {args.steps} frames in {t_full * 1e3:.1f} ms")

    # ---- pass B: per-frame GPU time (HIP event pair around each frame on the
    # library's stream; warp frames are graph launches), replaying the same
    # frames on fresh state, and PSNR against the full raycast
    def fresh():
        c = S.Context(0)
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
        if "tuned_warp" in res:
            t = res["tuned_warp"]
            A.apply(c, t)
        if not args.pretrace:
            c.set_pretrace(sets=0)
        for f in range(args.warmup):
            c.frame_warp(params[f])
        c.sync()
        return c

    # ---- pass A2: the same frames with the colorized image copied to host
    # memory after each (svo_read, blocking): the rate a caller that needs the
    # pixels on the host sees, PCIe copy included
    if args.legs == "all":
        wr = fresh()
        t0 = time.perf_counter()
        for f in range(args.warmup, total):
            wr.frame_warp(params[f])
            wr.image()
        t_readback = time.perf_counter() - t0
        wr.close()
        res["warp_fps_image_to_host"] = args.steps / t_readback
        log(f"[bench] warp + image to host: {args.steps} frames in {t_readback * 1e3:.1f} ms")
        # the same with svo_read_image_async into two pinned host buffers: each
        # copy overlaps the next frame (the library double-buffers the image)
        pins = [S.PinnedImage(rx, ry) for _ in range(2)]
        outs = [b.array for b in pins]
        wa = fresh()
        t0 = time.perf_counter()
        for f in range(args.warmup, total):
            wa.frame_warp(params[f])
            wa.read_image_async(outs[f & 1])
        wa.sync()
        t_async = time.perf_counter() - t0
        last = wa.image()
        ok = bool(np.array_equal(outs[(total - 1) & 1], last))
        wa.close()
        del outs
        for b in pins:
            b.close()
        res["warp_fps_image_to_host_async"] = args.steps / t_async
        log(f"[bench] warp + async image to pinned host memory: {args.steps} frames in {t_async * 1e3:.1f} ms "
            f"(last copy equals the device image: {ok})")
        if not ok:
            raise SystemExit("async image copy differs from the device image")

    # the same frames back to back without the pretrace (graph launches)
    wn = fresh()
    wn.set_pretrace(sets=0)
    t0 = time.perf_counter()
    for f in range(args.warmup, total):
        wn.frame_warp(params[f])
    wn.sync()
    res["warp_fps_no_pretrace"] = args.steps / (time.perf_counter() - t0)
    wn.close()
    log(f"[bench] warp without pretrace: {res['warp_fps_no_pretrace']:.1f} frames/s")

    warp2 = fresh()
    warp2.enable_timing(mode=2)
    full.enable_timing(mode=2)
    frame_ms, full_ms, psnrs, maxabs, same = [], [], [], [], []
    for f in range(args.warmup, total):
        warp2.frame_warp(params[f])
        warp2.sync()
        frame_ms.append(warp2.frame_time())
        if args.legs != "all":
            continue
        full.frame_full(params[f])
        full.sync()
        full_ms.append(full.frame_time())
        if (f - args.warmup) % max(1, args.steps // 10) == 0:
            a, b = S.image_rgb(warp2.image()), S.image_rgb(full.image())
            psnrs.append(S.psnr(a, b))
            d = np.abs(a.astype(np.int16) - b.astype(np.int16))
            maxabs.append(int(d.max()))
            same.append(float((d.max(axis=-1) == 0).mean()))
    warp2.close()
    # ---- pass C: per-stage times (an event before each stage, direct
    # launches), same frames on fresh state; the ray stage is the roofline's
    # kernel time
    warp3 = fresh()
    warp3.enable_timing(mode=1)
    full.enable_timing(mode=1)
    stage_acc, ray_ms_full = {}, []
    for f in range(args.warmup, total):
        warp3.frame_warp(params[f])
        warp3.sync()
        for k, v in warp3.stage_times().items():
            stage_acc.setdefault(k, []).append(v)
        if args.legs == "all" and f < args.warmup + 20:
            full.frame_full(params[f])
            full.sync()
            ray_ms_full.append(full.stage_times()["rays"])
    warp3.close()
    rays_per_frame = rays_warp / args.steps
    stage_mean = {k: float(np.mean(v)) for k, v in stage_acc.items()}
    res.update(
        warp_fps=args.steps / t_warp, full_fps=args.steps / t_full,
        warp_ms=frame_ms, full_ms=full_ms, stage_mean=stage_mean,
        rays_per_frame=rays_per_frame, rays_full_per_frame=rays_full / args.steps,
        mrays_warp=rays_warp / t_warp / 1e6, mrays_full=rays_full / t_full / 1e6,
        ray_ms_full=float(np.mean(ray_ms_full)) if ray_ms_full else float("nan"), psnr=psnrs, maxabs=maxabs,
        same=same, depth=depth, words=len(words),
        rx=rx, ry=ry)
    # ---- learned fill (gate + BF16 CNN) with the committed weights
    if args.learned and args.legs == "all":
        from svo_amd import learned as L
        if os.path.exists(L.WEIGHTS):
            cnn = S.Cnn(0)
            L.load(cnn, L.WEIGHTS)
            lw = S.Context(0)
            lw.upload_octree(words, root, depth)
            lw.resize(rx, ry)
            if args.tune:  # learned frames trace ~2x the rays of warp frames: their own knobs
                best, ms, _ = A.tune(lw, tp, warmup=10, mode="learned", cnn=cnn, max_cnn_holes=args.max_cnn_holes)
                res["tuned_learned"] = dict(best, ms=round(ms, 4))
                log(f"[bench] autotune learned: {best} -> {ms:.4f} ms/frame")
            for f in range(args.warmup):
                lw.frame_warp_learned(cnn, params[f], args.max_cnn_holes)
            lw.sync()
            t0 = time.perf_counter()
            for f in range(args.warmup, total):
                lw.frame_warp_learned(cnn, params[f], args.max_cnn_holes)
            lw.sync()
            res["learned_fps"] = args.steps / (time.perf_counter() - t0)
            lw.close()
            # per-stage times of the same frames on fresh state: the CNN stage
            # is the fused tile kernel; its work is the gate's CNN tile count
            lw = fresh()
            if "tuned_learned" in res:
                A.apply(lw, res["tuned_learned"])
            lw.enable_timing(mode=1)
            cnn_ms, tiles, gate_rays, lstage = [], [], [], {}
            for f in range(args.warmup, min(total, args.warmup + 40)):
                lw.frame_warp_learned(cnn, params[f], args.max_cnn_holes)
                lw.sync()
                st = lw.stage_times()
                for k, v in st.items():
                    lstage.setdefault(k, []).append(v)
                cnn_ms.append(st["cnn"])
                t_, g_ = lw.gate_counts()
                tiles.append(t_)
                gate_rays.append(g_)
            lw.close()
            res["learned_stage_ms"] = {k: round(float(np.mean(v)), 4) for k, v in lstage.items() if np.mean(v) > 0}
            res["cnn_ms"] = float(np.mean(cnn_ms))
            res["cnn_tiles"] = float(np.mean(tiles))
            res["gate_rays"] = float(np.mean(gate_rays))
            allp, holes = L.evaluate(cnn, name, seed_off=1000, n_frames=40, max_cnn_holes=args.max_cnn_holes)
            res["psnr_modes"] = {k: round(v, 2) for k, v in allp.items()}
            res["psnr_hole_pixels"] = {k: round(v, 2) for k, v in holes.items()}
            cnn.close()
        else:
            log(f"[bench] no trained weights at {L.WEIGHTS}: learned leg skipped")
    for c in (warp, full):
        c.close()
    return res


GATE_DEPTHS = [0.0, 4096.0, 1024.0, 256.0, 128.0, 96.0, 64.0]


def run_gate_curve(args, name, tuned=None):
    """Learned frames (gate + fused CNN, committed weights) along the gate's
    depth rule: for each depth limit (0 = off; world units) the back-to-back
    frames/s over the timed frames and the PSNR against full raycasting on a
    held-out path, with the share of pixels sent back to raycasting."""
    import svo_amd as S
    from svo_amd import autotune as A
    from svo_amd import learned as L
    from svo_amd import workloads as W

    if not os.path.exists(L.WEIGHTS):
        return None
    w = W.WORKLOADS[name]
    words, root, depth, _ = W.octree(name)
    rx, ry = w["res"]
    rows = W.path(name, args.warmup + args.steps)
    params = [W.frame_params(name, rows, f) for f in range(args.warmup + args.steps)]
    cnn = S.Cnn(0)
    L.load(cnn, L.WEIGHTS)
    out = []
    for d in GATE_DEPTHS:
        c = S.Context(0)
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
        if tuned:
            A.apply(c, tuned)
        c.set_gate_depth(d)
        for f in range(args.warmup):
            c.frame_warp_learned(cnn, params[f], args.max_cnn_holes)
        c.sync()
        t0 = time.perf_counter()
        for f in range(args.warmup, args.warmup + args.steps):
            c.frame_warp_learned(cnn, params[f], args.max_cnn_holes)
        c.sync()
        fps = args.steps / (time.perf_counter() - t0)
        gate_rays = c.gate_counts()[1]
        c.close()
        allp, holes, seq = L.evaluate(cnn, name, seed_off=1000, n_frames=24, skip=8,
                                      max_cnn_holes=args.max_cnn_holes, gate_depth=d, detail=True)
        db = lambda v: None if v is None or not np.isfinite(v) else round(float(v), 2)
        row = {"gate_depth": d, "fps": round(fps, 2), "psnr_learned_db": db(allp["learned"]),
               "psnr_gate_db": db(allp["gate"]), "psnr_plain_db": db(allp["plain"]),
               "seq_psnr_learned_db": db(seq["learned"]["seq_db"]), "seq_psnr_gate_db": db(seq["gate"]["seq_db"]),
               "identical_frames_learned": seq["learned"]["identical_frames"],
               "eval_frames": seq["learned"]["frames"], "gate_rays_last_frame": int(gate_rays)}
        log(f"[bench] gate curve {name}: {row}")
        out.append(row)
    cnn.close()
    return out


def run_stress(args, name="cave_4096", n_timed=190):
    """Divergence stress: the depth-12 cave over the whole 4096^3 volume with
    a fast camera (up to 32 voxels and 5 deg yaw per frame).  Scene generated and octree built on the
    GPU; warp frames (throughput and per-frame events, hole fraction read per
    frame) and full frames."""
    import svo_amd as S
    from svo_amd import workloads as W

    w = W.WORKLOADS[name]
    t = time.perf_counter()
    xyz, rgba = W.voxels(name)
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    words, root, depth, nvox = W.octree(name)
    t_build = time.perf_counter() - t
    log(f"[bench] {name}: {nvox} voxels generated in {t_gen:.2f}s, {len(words)} words built in {t_build:.2f}s")
    rx, ry = w["res"]
    total = 10 + n_timed
    rows = W.path(name, total)
    params = [W.frame_params(name, rows, f) for f in range(total)]
    warp = S.Context(0)
    full = S.Context(0)
    for c in (warp, full):
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
    tuned = None
    if args.tune and args.pretrace:
        # the pretrace only (off, its sets, mode, hardware queues and split
        # hand-out), on a held-out path: the cave's holes are a fifth of the
        # screen and its longest rays run ~770 trips, so tracing ahead may
        # not pay here
        from svo_amd import autotune as A
        trows = W.path(name, 50, seed_off=TUNE_SEED_OFF)
        tp = [W.frame_params(name, trows, f) for f in range(50)]
        best, ms, _ = A.tune(warp, tp, warmup=10, mode="warp",
                             space={"pre_sets": [4, 3, 0], "pre_pred": [0, 2], "pre_qskip": [1, 3, 0, 2, 4],
                                    "pre_split": [256, 0, 512]})
        tuned = dict(best, ms=round(ms, 4))
        log(f"[bench] stress autotune: {tuned}")
    for f in range(10):
        warp.frame_warp(params[f])
    warp.sync()
    warp.ray_stats(reset=True)
    warp.pretrace_stats(reset=True)
    t0 = time.perf_counter()
    for f in range(10, total):
        warp.frame_warp(params[f])
    warp.sync()
    t_warp = time.perf_counter() - t0
    rays, _ = warp.ray_stats(reset=True)
    pre_rays, merged, _ = warp.pretrace_stats(reset=True)
    rays += merged  # the frames' hole and tile rays
    # per-frame GPU time and hole fraction on fresh state
    w2 = S.Context(0)
    w2.upload_octree(words, root, depth)
    w2.resize(rx, ry)
    for f in range(10):
        w2.frame_warp(params[f])
    w2.sync()
    w2.enable_timing(mode=2)
    ms, holes = [], []
    for f in range(10, total):
        w2.frame_warp(params[f])
        w2.sync()
        ms.append(w2.frame_time())
        holes.append(w2.hole_count() / (rx * ry))
    full.enable_timing(mode=2)
    full.ray_stats(reset=True)
    fm = []
    for f in range(10, total):
        full.frame_full(params[f])
        full.sync()
        fm.append(full.frame_time())
    _, longest = full.ray_stats(reset=True)  # every pixel's ray of the timed frames
    for c in (warp, full, w2):
        c.close()
    return {"workload": f"{name}: {rx}x{ry}, SVO depth {depth}, 3-D noise cave shell over the whole volume, up to "
                        f"{w['vox']:.0f} voxels and {w['yaw']:.0f} deg yaw per frame",
            "voxels": int(nvox), "octree_words": int(len(words)), "scene_gpu_s": round(t_gen, 2),
            "build_gpu_s": round(t_build, 2), "warp_fps": round(n_timed / t_warp, 2),
            "warp_p50_ms": round(pct(ms, 50), 4), "warp_p99_ms": round(pct(ms, 99), 4),
            "full_p50_ms": round(pct(fm, 50), 4), "full_fps_from_p50": round(1000.0 / pct(fm, 50), 2),
            "frames": total, "timed_frames": n_timed,
            "hole_list_fraction_mean": round(float(np.mean(holes)), 4),
            "hole_list_fraction_max": round(float(np.max(holes)), 4),
            "rays_per_warp_frame": round(rays / n_timed), "rays_traced_ahead_per_frame": round(pre_rays / n_timed),
            "longest_ray_trips": int(longest), "tuned_pretrace": tuned}


def run_batch(args):
    """Batched views: 24 seeded cameras of uneven resolution over the depth-11
    scene.  The same frames run once view by view (svo_frame_warp per
    context, back to back) and once as batched frames (one launch per stage
    over all views, one shared ray launch)."""
    import svo_amd as S
    from svo_amd import autotune as A
    from svo_amd import workloads as W

    views = W.batch_views()
    words, root, depth, _ = W.octree(views[0]["scene"])
    n_frames = views[0]["frames"]
    frames = []
    paths = [W.view_path(v) for v in views]
    for f in range(n_frames):
        frames.append([W.frame_params(v["scene"], paths[i], f) for i, v in enumerate(views)])
    ctxs = []
    for v in views:
        c = S.Context(0)
        c.upload_octree(words, root, depth)
        c.resize(*v["res"])
        c.set_pretrace(sets=0)  # view-by-view frames as the batch runs them: no pretrace
        ctxs.append(c)
    pixels = sum(v["res"][0] * v["res"][1] for v in views)
    warm = max(2, min(args.batch_warmup, n_frames - 1))
    out = dict(views=len(views), frames_per_view=n_frames, warmup_frames=warm,
               pixels_per_batch_frame=pixels,
               res_counts={f"{rx}x{ry}": sum(1 for v in views if v["res"] == (rx, ry)) for rx, ry in W.VIEW_RES})

    def seq_run():
        for ps in frames[:warm]:
            for c, p in zip(ctxs, ps):
                c.frame_warp(p)
        for c in ctxs:
            c.sync()
            c.ray_count(reset=True)
        t0 = time.perf_counter()
        for ps in frames[warm:]:
            for c, p in zip(ctxs, ps):
                c.frame_warp(p)
        for c in ctxs:
            c.sync()
        t = time.perf_counter() - t0
        return t, sum(c.ray_count(reset=True) for c in ctxs)

    batch = S.Batch(ctxs)

    def batch_run():
        for ps in frames[:warm]:
            batch.frame_warp(ps)
        batch.sync()
        ctxs[0].ray_count(reset=True)
        t0 = time.perf_counter()
        for ps in frames[warm:]:
            batch.frame_warp(ps)
        batch.sync()
        t = time.perf_counter() - t0
        return t, ctxs[0].ray_count(reset=True)

    if args.tune:  # on held-out view paths
        from svo_amd import autotune as A
        tpaths = [W.view_path(v, seed_off=TUNE_SEED_OFF) for v in views]
        tframes = [[W.frame_params(v["scene"], tpaths[i], f) for i, v in enumerate(views)]
                   for f in range(min(n_frames, 16))]
        best, ms, _ = A.tune_batch(batch, tframes, warmup=6)
        out["tuned"] = dict(best, ms=round(ms, 4))
        log(f"[bench] batch autotune: {best} -> {ms:.4f} ms/batch frame")
        for c in ctxs[1:]:  # the view-by-view run uses the same knobs (faster there than the defaults)
            A.apply(c, best)
    t_seq, rays_seq = seq_run()
    t_bat, rays_bat = batch_run()
    k = n_frames - warm
    # the same frames with the learned fill on (gate + fused CNN per view)
    from svo_amd import learned as L
    if args.learned and os.path.exists(L.WEIGHTS):
        cnn = S.Cnn(0)
        L.load(cnn, L.WEIGHTS)
        for ps in frames[:warm]:
            batch.frame_warp_learned(cnn, ps, args.max_cnn_holes)
        batch.sync()
        t0 = time.perf_counter()
        for ps in frames[warm:]:
            batch.frame_warp_learned(cnn, ps, args.max_cnn_holes)
        batch.sync()
        t_bl = time.perf_counter() - t0
        out["learned_batch_view_frames_per_s"] = round(k * len(views) / t_bl, 1)
        cnn.close()
    out.update(
        batch_frames_per_s=round(k / t_bat, 2), batch_view_frames_per_s=round(k * len(views) / t_bat, 1),
        sequential_view_frames_per_s=round(k * len(views) / t_seq, 1),
        speedup_batch_vs_sequential=round(t_seq / t_bat, 3),
        batch_ms_per_frame=round(t_bat / k * 1e3, 4),
        mrays_s_batch=round(rays_bat / t_bat / 1e6, 1), mrays_s_sequential=round(rays_seq / t_seq / 1e6, 1),
        mpixels_s_batch=round(pixels * k / t_bat / 1e6, 1))
    log(f"[bench] batch: {out}")
    batch.close()
    for c in ctxs:
        c.close()
    return out


def host_cpu():
    """The host the CPU baseline runs on: logical CPUs (nproc), the CPUs this
    process may run on, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return dict(nproc=os.cpu_count(), allowed=allowed, model=model)


def fetch_words_per_ray(name, first, last, threads):
    """F-bar of the roofline: mean octree words a ray of the warp frames
    first..last-1 reads (the oracle replays the path from frame 0 and counts
    them on exactly those frames, the ones the GPU timing covers)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from svo_amd import workloads as W

    w = W.WORKLOADS[name]
    words, root, depth, _ = W.octree(name)
    rows = W.path(name)
    fr = O.Frame(*w["res"], words, root, depth, nthreads=threads)
    for f in range(first):
        fr.frame_warp(O.params_from(W.frame_params(name, rows, f)))
    st0 = fr.stats
    for f in range(first, last):
        fr.frame_warp(O.params_from(W.frame_params(name, rows, f)))
    st1 = fr.stats
    fr.close()
    return (st1["fetch_words"] - st0["fetch_words"]) / max(1, st1["rays"] - st0["rays"])


def run_cpu_baseline(name, n_warp_frames, n_full_frames, threads):
    """Oracle (the CPU restatement, kind "port") on a bounded sample of the
    same workload: frames 0-1 (full raycasts through the hole path) untimed,
    then n_warp_frames warp frames timed; plus n_full_frames full frames."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from svo_amd import workloads as W

    w = W.WORKLOADS[name]
    words, root, depth, _ = W.octree(name)
    rx, ry = w["res"]
    rows = W.path(name)
    fr = O.Frame(rx, ry, words, root, depth, nthreads=threads)
    for f in range(2):
        fr.frame_warp(O.params_from(W.frame_params(name, rows, f)))
    t0 = time.perf_counter()
    for f in range(2, 2 + n_warp_frames):
        fr.frame_warp(O.params_from(W.frame_params(name, rows, f)))
    t_warp = time.perf_counter() - t0
    fr2 = O.Frame(rx, ry, words, root, depth, nthreads=threads)
    t0 = time.perf_counter()
    for f in range(2, 2 + n_full_frames):
        fr2.frame_full(O.params_from(W.frame_params(name, rows, f)))
    t_full = time.perf_counter() - t0
    fr.close()
    fr2.close()
    return dict(warp_fps=n_warp_frames / t_warp, full_fps=n_full_frames / t_full, full_frames=n_full_frames,
                t=t_warp + t_full, frames=n_warp_frames)


LONE_FRAME, LONE_PIXEL, LONE_TRIPS = 25, (216, 827), 3450  # flythrough_2048: the longest ray of frame 25


def run_lone_ray(args):
    """The latency of one long ray: the longest primary ray of frame 25 of
    the flythrough (3450 loop trips) launched alone, then its 8x8 block (64
    long rays in one wave), 20 launches each, one persistent workgroup per
    CU.  Wall time per launch + sync here; the kernel time comes from a
    rocprofv3 kernel trace of this run."""
    import svo_amd as S
    from svo_amd import workloads as W

    name = "flythrough_2048"
    words, root, depth, _ = W.octree(name)
    rx, ry = W.WORKLOADS[name]["res"]
    p = W.frame_params(name, W.path(name, LONE_FRAME + 1), LONE_FRAME)
    out = {}
    with S.Context(0) as c:
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
        c.set_tuning(-1, 1, -1, 0)
        x, y = LONE_PIXEL
        for label, rect in (("lone_ray", (x, y, 1, 1)), ("block_8x8", (x & ~7, y & ~7, 8, 8))):
            for _ in range(3):
                c.raycast_rect(p, *rect)
            c.sync()
            c.ray_stats(reset=True)
            ts = []
            for _ in range(20):
                t0 = time.perf_counter()
                c.raycast_rect(p, *rect)
                c.sync()
                ts.append((time.perf_counter() - t0) * 1e3)
            n, longest = c.ray_stats(reset=True)
            out[label] = {"rect": list(rect), "longest_trips": int(longest), "wall_ms_p50": round(pct(ts, 50), 4)}
            log(f"[bench] {label}: {out[label]}")
    if out["lone_ray"]["longest_trips"] != LONE_TRIPS:
        raise SystemExit(f"lone ray: {out['lone_ray']['longest_trips']} trips, expected {LONE_TRIPS}")
    print(json.dumps({"lone_ray": out}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="flythrough_2048")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-frames", type=int, default=280, help="warp frames in the CPU sample")
    ap.add_argument("--cpu-full-frames", type=int, default=48, help="full frames in the CPU sample")
    ap.add_argument("--no-tune", dest="tune", action="store_false", help="skip the ray-kernel autotuner")
    ap.add_argument("--no-pretrace", dest="pretrace", action="store_false",
                    help="warp frames without a pretrace (single stream; for counter runs, which serialise dispatches)")
    ap.add_argument("--no-learned", dest="learned", action="store_false", help="skip the learned-fill leg")
    ap.add_argument("--max-cnn-holes", type=int, default=64)
    ap.add_argument("--no-batch", dest="batch", action="store_false", help="skip the batched-views leg")
    ap.add_argument("--no-gate-curve", dest="gate_curve", action="store_false",
                    help="skip the learned-fill frames/s vs PSNR sweep of the gate's depth rule")
    ap.add_argument("--no-stress", dest="stress", action="store_false",
                    help="skip the divergence-stress leg (depth-12 cave)")
    ap.add_argument("--batch-warmup", type=int, default=8)
    ap.add_argument("--legs", default="all", choices=["all", "warp", "lone"],
                    help="warp: only the warp-frame legs (for profiling the warp kernels); lone: only the "
                         "single-long-ray launches")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); sharding across GPUs is a "
              "design, not built.", file=sys.stderr)
        sys.exit(2)
    if args.warmup < 2:
        raise SystemExit("--warmup must be >= 2 (frames 0-1 are full clears)")
    if args.legs == "lone":
        run_lone_ray(args)
        return

    g = run_gpu(args, args.workload)
    curve = None
    if args.learned and args.legs == "all" and args.gate_curve:
        tl = g.get("tuned_learned")
        curve = {"flythrough_2048": run_gate_curve(args, args.workload, tl),
                 "flythrough_2048_flat": run_gate_curve(args, "flythrough_2048_flat", tl)}
    batch = run_batch(args) if (args.batch and args.legs == "all") else None
    stress = run_stress(args) if (args.stress and args.legs == "all") else None
    if stress:
        log(f"[bench] divergence stress: {stress}")
    # dominant kernel: ray traversal in the warp frames
    ray_ms = g["stage_mean"].get("rays", float("nan"))
    cpu = None
    host = host_cpu()
    # the CPU legs use the CPUs this process may run on, at most the GPU
    # box's CPU share of 16 per GPU
    threads = max(1, min(CPU_SHARE, host["allowed"]))
    fbar = fetch_words_per_ray(args.workload, args.warmup, args.warmup + args.steps, threads)
    log(f"[bench] F-bar over warp frames {args.warmup}..{args.warmup + args.steps - 1}: {fbar:.2f} words/ray")
    if not args.no_cpu:
        cpu = run_cpu_baseline(args.workload, args.cpu_frames, args.cpu_full_frames, threads)
        log(f"[bench] cpu baseline: {cpu} on {host}")
    roof = None
    if fbar is not None and ray_ms == ray_ms:
        # algorithmic bytes per launch: per ray 4 B id + 4 B word + 12 B xyz + 4 B per octree word
        bytes_launch = g["rays_per_frame"] * (20.0 + 4.0 * fbar)
        achieved = bytes_launch / (ray_ms * 1e-3) / 1e9
        roof = {"kernel": "k_rays", "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "per_launch_bytes": round(bytes_launch), "avg_launch_ms": round(ray_ms, 4),
                "fetch_words_per_ray": round(fbar, 2),
                "fetch_words_frames": [args.warmup, args.warmup + args.steps - 1]}
        # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
        # runs of this workload, scaled by a copy kernel of known size
        tpath = os.path.join(ROOT, "profiles", "k_rays_traffic.json")
        if os.path.exists(tpath) and args.workload == "flythrough_2048":
            with open(tpath) as fh:
                t = json.load(fh)
            roof["traffic"] = t["traffic_bytes"]
            roof["traffic_source"] = t["source"]
    wl = g
    out = {
        "metric": METRIC,
        "value": round(g["warp_fps"], 2),
        "unit": "frames/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000.0 / g["warp_fps"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {wl['rx']}x{wl['ry']}, SVO depth {wl['depth']}, "
                               "warp+fill frames along the seeded flythrough",
                   "res": [wl["rx"], wl["ry"]], "octree_words": wl["words"]},
        # per-frame GPU time of frames issued one at a time (host waits for
        # each; no pretrace then), and the time between the completions of
        # successive frames issued back to back (the timed frames)
        "warp_p50_ms": round(pct(g["warp_ms"], 50), 4), "warp_p99_ms": round(pct(g["warp_ms"], 99), 4),
        "warp_interval_p50_ms": round(pct(g["warp_intervals"], 50), 4),
        "warp_interval_p99_ms": round(pct(g["warp_intervals"], 99), 4),
        "warp_fps_no_pretrace": round(g["warp_fps_no_pretrace"], 2),
        "pretrace": g["pretrace"],
        "warp_fps_image_to_host": round(g["warp_fps_image_to_host"], 2) if "warp_fps_image_to_host" in g else None,
        "warp_fps_image_to_host_async": (round(g["warp_fps_image_to_host_async"], 2)
                                         if "warp_fps_image_to_host_async" in g else None),
        "full_fps": round(g["full_fps"], 2),
        "full_p50_ms": round(pct(g["full_ms"], 50), 4), "full_p99_ms": round(pct(g["full_ms"], 99), 4),
        "full_interval_p50_ms": num(pct(g.get("full_intervals", []), 50)),
        "full_interval_p99_ms": num(pct(g.get("full_intervals", []), 99)),
        "speedup_warp_vs_full": round(g["warp_fps"] / g["full_fps"], 3),
        "mrays_s_warp": round(g["mrays_warp"], 1), "mrays_s_full": round(g["mrays_full"], 1),
        "rays_per_warp_frame": round(g["rays_per_frame"]),
        "psnr_warp_vs_full_db": round(float(np.mean([p for p in g["psnr"] if p != float("inf")] or [99.0])), 2),
        "maxabs_warp_vs_full": max(g["maxabs"]) if g["maxabs"] else None,
        "identical_pixels_warp_vs_full": round(float(np.mean(g["same"])), 4) if g["same"] else None,
        "stage_ms": {k: round(v, 4) for k, v in g["stage_mean"].items()},
        "autotuned": {k: g[k] for k in ("tuned_warp", "tuned_full", "tuned_learned") if k in g},
        "learned_fill": None if "learned_fps" not in g else {
            "fps": round(g["learned_fps"], 2), "speedup_vs_full": round(g["learned_fps"] / g["full_fps"], 3),
            "psnr_vs_full_db": g["psnr_modes"], "psnr_hole_pixels_db": g["psnr_hole_pixels"],
            "max_cnn_holes_per_tile": args.max_cnn_holes,
            "stage_ms": g["learned_stage_ms"], "cnn_tiles_per_frame": round(g["cnn_tiles"]),
            "gate_rays_per_frame": round(g["gate_rays"]),
            # fused tile kernel, bf16 MFMA: useful FLOPs (real channels, the
            # regions the centre 16x16 needs) and FLOPs the MFMAs issue
            "roofline": {"kernel": "k_cnn_tiles", "bound": "mfma", "unit": "TFLOP/s",
                         "achieved": round(g["cnn_tiles"] * CNN_TILE_FLOPS / (g["cnn_ms"] * 1e-3) / 1e12, 2),
                         "issued": round(g["cnn_tiles"] * CNN_TILE_MFMA_FLOPS / (g["cnn_ms"] * 1e-3) / 1e12, 2),
                         "peak": BF16_PEAK_TFLOPS,
                         "frac": round(g["cnn_tiles"] * CNN_TILE_FLOPS / (g["cnn_ms"] * 1e-3) / 1e12
                                       / BF16_PEAK_TFLOPS, 5),
                         "avg_launch_ms": round(g["cnn_ms"], 4), "flops_per_tile": CNN_TILE_FLOPS},
            # frames/s vs PSNR along the gate's depth rule (tiles with a warped
            # pixel deeper than gate_depth world units raycast whole); the flat
            # scene is the same terrain without per-voxel colour jitter
            "gate_depth_curve": curve},
        "octree_build": g["octree_build"],
        "batched_views": batch,
        "divergence_stress": stress,
        "roofline": roof,
        "cpu_baseline": None if cpu is None else {
            "value": round(cpu["warp_fps"], 3), "unit": "frames/s", "cores": threads, "kind": "port",
            "host_nproc": host["nproc"], "host_cpus_allowed": host["allowed"], "cpu_model": host["model"],
            "full_fps": round(cpu["full_fps"], 3),
            "sample": f"{cpu['frames']} warp frames (after 2 clear frames) + {cpu['full_frames']} full "
                      f"frames of the same workload, oracle restatement, {threads} threads on "
                      f"{host['allowed']} allowed of {host['nproc']} host CPUs ({host['model']}), "
                      f"{cpu['t']:.1f} s"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
