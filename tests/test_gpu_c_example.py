"""The C host-loop example (sparse-voxel-octree-raycasting_amd/examples/
render_flythrough.c) on a GPU: a program that uses only include/svo.h and
libsvo.so renders the depth-8 flythrough, and its PNG holds exactly the
image the same frames give through the ctypes binding."""
import os
import subprocess

import numpy as np
import pytest

import svo_amd as S
from helpers import CONFIG1, params, path, read_png, scene_octree

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sparse-voxel-octree-raycasting_amd", "examples", "render_flythrough")


@pytest.mark.gpu
def test_c_example_matches_the_binding(tmp_path):
    frames = 6
    out = tmp_path / "c.png"
    print(f"c example: {frames} frames", flush=True)
    r = subprocess.run([EXE, str(out), str(frames)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rgb = read_png(out.read_bytes())

    words, root = scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])
    rx, ry = CONFIG1["res"]
    rows = path(CONFIG1, frames)
    c = S.Context(0)
    try:
        c.upload_octree(words, root, CONFIG1["depth"])
        c.resize(rx, ry)
        for f in range(frames):
            c.frame_warp(params(CONFIG1, rows, f))
        img = c.image()
    finally:
        c.close()
    want = np.stack([(img >> 16) & 255, (img >> 8) & 255, img & 255], axis=-1).astype(np.uint8)
    assert rgb.shape == want.shape
    assert np.array_equal(rgb, want)
