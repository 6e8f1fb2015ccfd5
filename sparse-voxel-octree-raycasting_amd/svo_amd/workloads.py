"""Seeded synthetic workloads: scenes, flythrough paths and per-frame camera
arguments.  Scene seed = 1000 + n, path seed = 2000 + n for workload n.

  flythrough_256:  depth 8 (256^3) 2-octave value-noise heightfield shell,
                   512x256, 64 frames, <= 2 voxels and <= 0.5 deg yaw per frame
  flythrough_2048: depth 11 (2048^3) 4-octave fractal terrain shell + 2000
                   sphere shells (radius 4..64 voxels), 1920x1024, 300 frames,
                   1..8 voxels and <= 1 deg yaw per frame
  cave_4096:       depth 12 (4096^3) 5-octave 3-D noise cave over the whole
                   volume (boundary voxels), 1920x1024, 200 frames, up to 32
                   voxels and 5 deg yaw per frame (divergence stress)
  flythrough_2048_flat: flythrough_2048's scene and path without the per-voxel
                   colour jitter (voxel colours by height only); the jitter
                   sets the reprojection error of warped pixels
  views_24:        24 independent cameras over the flythrough_2048 scene,
                   per-view resolution drawn from {320x240, 640x480,
                   1280x720, 1920x1024}, per-view motion up to 0.5..16
                   voxels and 0..3 deg yaw per frame, 120 frames per view
"""
from __future__ import annotations

import functools

import numpy as np

from . import build_octree, build_octree_gpu, camera, camera_path, scene, scene_gpu

WORKLOADS = {
    "flythrough_256": dict(kind=1, seed=1001, depth=8, spheres=0, res=(512, 256), frames=64,
                           vox=2.0, yaw=0.5, height=240.0, pitch=0.35, path_seed=2001),
    "flythrough_2048": dict(kind=2, seed=1002, depth=11, spheres=2000, res=(1920, 1024), frames=300,
                            vox=8.0, yaw=1.0, height=1400.0, pitch=0.30, path_seed=2002),
    "flythrough_2048_flat": dict(kind=5, seed=1002, depth=11, spheres=2000, res=(1920, 1024), frames=300,
                                 vox=8.0, yaw=1.0, height=1400.0, pitch=0.30, path_seed=2002),
    "cave_4096": dict(kind=4, seed=1003, depth=12, spheres=0, res=(1920, 1024), frames=200,
                      vox=32.0, yaw=5.0, height=2048.0, pitch=0.05, path_seed=2003),
}


@functools.lru_cache(maxsize=2)
def voxels(name):
    """(xyz, rgba) voxel stream of a workload's scene (the cave on the GPU:
    the host generator needs hours at depth 12; both give the same stream)."""
    w = WORKLOADS[name]
    if w["kind"] in (3, 4):
        return scene_gpu(w["kind"], w["seed"], w["depth"])
    return scene(w["kind"], w["seed"], w["depth"], w["spheres"])


@functools.lru_cache(maxsize=4)
def octree(name):
    """(words, root_ref, depth, n_voxels) of a workload's scene (the cave's
    octree is built on the GPU; the host builder gives the same words)."""
    w = WORKLOADS[name]
    xyz, rgba = voxels(name)
    build = build_octree_gpu if w["kind"] in (3, 4) else build_octree
    words, root = build(xyz, rgba, w["depth"])
    return words, root, w["depth"], len(rgba)


def path(name, n_frames=None, seed_off=0):
    """The workload's seeded camera path; seed_off != 0 gives a held-out path
    of the same motion bounds (used to tune without seeing the timed frames)."""
    w = WORKLOADS[name]
    return camera_path(w["path_seed"] + seed_off, w["depth"], n_frames or w["frames"], w["vox"], w["yaw"],
                       w["height"], w["pitch"])


def frame_params(name, rows, f, frame_index=None):
    """FrameParams of path row f (frame counter frame_index, default f)."""
    w = WORKLOADS[name]
    yaw, pitch, roll, x, y, z = [float(v) for v in rows[f]]
    p, _ = camera(yaw, pitch, roll, (x, y, z), w["depth"], f if frame_index is None else frame_index)
    return p


VIEW_RES = [(320, 240), (640, 480), (1280, 720), (1920, 1024)]


def batch_views(n_views=24, seed=5005, frames=120, scene="flythrough_2048"):
    """Per-view dicts (res, vox, yaw, path_seed) of the batched-views
    workload, drawn from seed."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_views):
        res = VIEW_RES[int(rng.integers(len(VIEW_RES)))]
        out.append(dict(scene=scene, res=res, vox=float(rng.uniform(0.5, 16.0)),
                        yaw=float(rng.uniform(0.0, 3.0)), path_seed=seed + 1 + i, frames=frames))
    return out


def view_path(view, seed_off=0):
    w = WORKLOADS[view["scene"]]
    return camera_path(view["path_seed"] + seed_off, w["depth"], view["frames"], view["vox"], view["yaw"],
                       w["height"], w["pitch"])
