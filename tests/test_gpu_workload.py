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
    rays (traced in the frame or taken from its pretrace) = holes + tile
    pixels."""
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
            c.pretrace_stats(reset=True)
            c.frame_warp(p)
            c.sync()
            if f % 8 != 5:
                continue
            ids = c.ids()
            H = int(ids[0])
            _, _, ex, ey = S.tile_schedule(rx, ry, f)
            assert c.ray_count() + c.pretrace_stats()[1] == H + ex * ey
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


@pytest.mark.parametrize("res", [(8192, 8192), (8192, 16), (16, 8192)])
def test_maximum_resolution_sampled_vs_oracle(big, res):
    """The largest (and most lopsided) frames svo_resize accepts: a full frame
    checked against the oracle's single-ray trace on 1500 seeded pixels, then
    warp frames whose id list matches plane 0's hole quads and whose graph
    replay matches direct launches on the same pixels."""
    words, root, depth = big
    rx, ry = res
    rows = W.path(NAME, 4)
    rng = np.random.default_rng(rx * 7 + ry)
    xs = rng.integers(0, rx, 1500)
    ys = rng.integers(0, ry, 1500)
    with S.Context(0) as c:
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
        p = W.frame_params(NAME, rows, 3)
        print(f"full frame {rx}x{ry}", flush=True)
        c.frame_full(p)
        c.sync()
        scr = c.screen(0).reshape(ry, rx)
        crd = c.coords(0).reshape(ry, rx, 4)
        op = O.params_from(p)
        for x, y in zip(xs, ys):
            w, xyz, _ = O.trace_pixel(words, root, depth, rx, ry, op, int(x), int(y))
            assert int(scr[y, x]) == w, (x, y)
            assert_equal_f32_bits(crd[y, x, :3], np.array(xyz, np.float32), f"pixel {x},{y}")
        # warp frames: graph (frames >= 2) on c, direct launches on d
        with S.Context(0) as d:
            d.upload_octree(words, root, depth)
            d.resize(rx, ry)
            d.set_graph(False)
            for f in range(4):
                p = W.frame_params(NAME, rows, f)
                print(f"warp frame {f} {rx}x{ry}", flush=True)
                c.frame_warp(p)
                d.frame_warp(p)
                c.sync()
                d.sync()
                a, b = c.screen(0).reshape(-1), d.screen(0).reshape(-1)
                sel = ys.astype(np.int64) * rx + xs
                assert np.array_equal(a[sel], b[sel])
                assert int(c.ids()[0]) == int(d.ids()[0])


def test_ray_stats_match_the_oracle(big):
    """svo_ray_stats: per frame, the rays traced and the loop trips of the
    longest of them equal the oracle's count of rays and its longest ray
    (warp frames after the clears, then a full frame; no pretrace, so every
    ray of the frame is traced in it)."""
    words, root, depth = big
    rx, ry = W.WORKLOADS[NAME]["res"]
    rows = W.path(NAME, 10)
    o = O.Frame(rx, ry, words, root, depth, nthreads=16)
    try:
        with S.Context(0) as c:
            c.upload_octree(words, root, depth)
            c.resize(rx, ry)
            c.set_pretrace(0)
            for f in range(9):
                p = W.frame_params(NAME, rows, f)
                print(f"ray stats frame {f}", flush=True)
                c.ray_stats(reset=True)
                o.s.st = O.Stats()
                if f < 8:
                    c.frame_warp(p)
                    o.frame_warp(O.params_from(p))
                else:
                    c.frame_full(p)
                    o.frame_full(O.params_from(p))
                rays, longest = c.ray_stats()
                st = o.stats
                assert (rays, longest) == (st["rays"], st["max_iters"]), f"frame {f}"
                assert longest > 100
    finally:
        o.close()
