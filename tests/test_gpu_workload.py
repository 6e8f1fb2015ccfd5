"""The headline workload (1920x1024, depth-11 terrain + 2000 spheres) on the
GPU vs the oracle: the first warp frames bit-exact, and size-independent
properties over a longer run."""
import numpy as np
import pytest

import oracle as O
import svo_amd as S
from svo_amd import workloads as W
from helpers import assert_equal_f32_bits, assert_equal_u32

pytestmark = pytest.mark.gpu
NAME = "flythrough_2048"


@pytest.fixture(scope="module")
def big():
    words, root, depth, _ = W.octree(NAME)
    return words, root, depth


def test_workload_frames_bitexact(big):
    words, root, depth = big
    rx, ry = W.WORKLOADS[NAME]["res"]
    rows = W.path(NAME, 8)
    o = O.Frame(rx, ry, words, root, depth, nthreads=16)
    with S.Context(0) as c:
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
        for f in range(5):
            p = W.frame_params(NAME, rows, f)
            print(f"workload frame {f}", flush=True)
            c.frame_warp(p)
            o.frame_warp(O.params_from(p))
            c.sync()
            assert_equal_u32(c.screen(), o.screen, f"frame {f} screen")
            assert_equal_f32_bits(c.coords(), o.coords, f"frame {f} coords")
            assert_equal_u32(c.ids(), o.ids, f"frame {f} ids")
            assert_equal_u32(c.image(), o.image, f"frame {f} image")
    o.close()


def test_workload_properties(big):
    """Over 40 frames: the id list is the ascending-block-order list of hole
    quads of plane 0 (4 ids per quad, counts = 4 x quads), and the frame's
    rays = holes + tile pixels."""
    words, root, depth = big
    rx, ry = W.WORKLOADS[NAME]["res"]
    rows = W.path(NAME, 40)
    S_ = (rx // 16) * (ry // 16)
    with S.Context(0) as c:
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
        for f in range(40):
            p = W.frame_params(NAME, rows, f)
            c.ray_count(reset=True)
            c.frame_warp(p)
            c.sync()
            if f % 8 != 5:
                continue
            ids = c.ids()
            H = int(ids[0])
            _, _, ex, ey = S.tile_schedule(rx, ry, f)
            assert c.ray_count() == H + ex * ey
            offs = ids[S_:2 * S_]
            assert offs[0] == 0 and np.all(np.diff(offs.astype(np.int64)) >= 0)
            assert H % 4 == 0 and H <= rx * ry
            lst = ids[2 * S_:2 * S_ + H].reshape(-1, 4)
            x, y = lst[:, 0] & 0xFFFF, lst[:, 0] >> 16
            assert np.all(x % 2 == 0) and np.all(y % 2 == 0)
            assert np.all(lst[:, 1] == lst[:, 0] + 1)
            assert np.all(lst[:, 2] == lst[:, 0] + 1 + (1 << 16))
            assert np.all(lst[:, 3] == lst[:, 0] + (1 << 16))
            blk = (y // 16) * (rx // 16) + x // 16
            assert np.all(np.diff(blk.astype(np.int64)) >= 0)
