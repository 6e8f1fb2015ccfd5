"""Seeded synthetic workloads: scenes, flythrough paths and per-frame camera
arguments.  Scene seed = 1000 + n, path seed = 2000 + n for workload n.

  flythrough_256:  depth 8 (256^3) 2-octave value-noise heightfield shell,
                   512x256, 64 frames, <= 2 voxels and <= 0.5 deg yaw per frame
  flythrough_2048: depth 11 (2048^3) 4-octave fractal terrain shell + 2000
                   sphere shells (radius 4..64 voxels), 1920x1024, 300 frames,
                   1..8 voxels and <= 1 deg yaw per frame
"""
from __future__ import annotations

import functools

from . import build_octree, camera, camera_path, scene

WORKLOADS = {
    "flythrough_256": dict(kind=1, seed=1001, depth=8, spheres=0, res=(512, 256), frames=64,
                           vox=2.0, yaw=0.5, height=240.0, pitch=0.35, path_seed=2001),
    "flythrough_2048": dict(kind=2, seed=1002, depth=11, spheres=2000, res=(1920, 1024), frames=300,
                            vox=8.0, yaw=1.0, height=1400.0, pitch=0.30, path_seed=2002),
}


@functools.lru_cache(maxsize=4)
def octree(name):
    """(words, root_ref, depth, n_voxels) of a workload's scene."""
    w = WORKLOADS[name]
    xyz, rgba = scene(w["kind"], w["seed"], w["depth"], w["spheres"])
    words, root = build_octree(xyz, rgba, w["depth"])
    return words, root, w["depth"], len(rgba)


def path(name, n_frames=None):
    w = WORKLOADS[name]
    return camera_path(w["path_seed"], w["depth"], n_frames or w["frames"], w["vox"], w["yaw"],
                       w["height"], w["pitch"])


def frame_params(name, rows, f, frame_index=None):
    """FrameParams of path row f (frame counter frame_index, default f)."""
    w = WORKLOADS[name]
    yaw, pitch, roll, x, y, z = [float(v) for v in rows[f]]
    p, _ = camera(yaw, pitch, roll, (x, y, z), w["depth"], f if frame_index is None else frame_index)
    return p
