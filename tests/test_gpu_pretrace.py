"""svo_set_pretrace: warp frames whose refresh rectangle and long-ray blocks
(or predicted hole quads) are traced ahead on other streams equal frames
traced entirely in the frame (and the oracle), for every set count, block
threshold, prediction depth and launch path; the frame rays that take a
result traced ahead are counted, and every traced or merged ray is one of
the frame's hole and rectangle rays."""
import numpy as np
import pytest

import oracle as O
import svo_amd as S
from svo_amd import workloads as W
from helpers import CONFIG1, assert_equal_f32_bits, assert_equal_u32, params, path, scene_octree

pytestmark = pytest.mark.gpu
FLY = "flythrough_2048"


@pytest.fixture(scope="module")
def scene1():
    return scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])


def _same(a, b, what):
    assert_equal_u32(a.screen(), b.screen(), f"{what} screen")
    assert_equal_f32_bits(a.coords(), b.coords(), f"{what} coords")
    assert_equal_u32(a.ids(), b.ids(), f"{what} ids")
    assert_equal_u32(a.image(), b.image(), f"{what} image")


@pytest.mark.parametrize("sets,long_trips,graph,wg,split,lanes", [
    (3, 512, True, 0, 0, 8), (2, 64, True, 0, 0, 8), (8, 1, True, 0, 0, 8), (3, 1 << 22, True, 0, 0, 8),
    (3, 128, False, 0, 0, 8), (5, 256, False, 1, 1024, 8), (4, 256, True, 2, 512, 1), (3, 64, False, 0, 64, 64),
    (6, 16, False, 1, 16, 4)])
def test_pretrace_frames_equal_plain_frames(scene1, sets, long_trips, graph, wg, split, lanes):
    words, root = scene1
    rx, ry = CONFIG1["res"]
    n = 24
    pr = path(CONFIG1, n)
    S_ = (rx // 16) * (ry // 16)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, CONFIG1["depth"])
            c.resize(rx, ry)
            c.set_graph(graph)
        # long blocks (predict 0), every frame, synchronised or not
        a.set_pretrace(sets, long_trips, wg, split, lanes, only_queued=0, predict=0)
        cfg = a.get_pretrace()
        assert (cfg.sets, cfg.long_trips, cfg.wg_per_cu, cfg.split_trips, cfg.split_lanes, cfg.only_queued) == \
            (sets, long_trips, wg, split, lanes, 0)
        b.set_pretrace(0)
        merged_total = 0
        for f in range(n):
            print(f"pretrace sets={sets} long={long_trips} graph={graph} wg={wg} split={split}/{lanes} frame {f}",
                  flush=True)
            p = params(CONFIG1, pr, f)
            a.ray_stats(reset=True)
            a.pretrace_stats(reset=True)
            a.frame_warp(p)
            b.frame_warp(p)
            a.sync()
            rays, _ = a.ray_stats()
            pre_rays, merged, _ = a.pretrace_stats()
            H = int(b.ids()[0])
            _, _, ex, ey = S.tile_schedule(rx, ry, f)
            assert rays + merged == H + ex * ey, f"frame {f}"
            if f >= 2:
                assert pre_rays >= ex * ey and merged >= ex * ey
                if long_trips == 1:  # every block is long: the frame traces nothing itself
                    assert rays == 0 and pre_rays == rx * ry
            else:
                assert pre_rays == 0 and merged == 0
            merged_total += merged
            if f % 4 == 3 or f == n - 1:
                _same(a, b, f"frame {f}")
        assert merged_total > 0
    finally:
        a.close()
        b.close()


def test_pretrace_flythrough_vs_oracle():
    """The headline workload's frames 0..29 with the default pretrace (the
    bench's frames 5..24 among them) against the oracle, plane 0, the id
    list and the image every frame; the long blocks carry most of the frame's
    longest rays."""
    words, root, depth, _ = W.octree(FLY)
    rx, ry = W.WORKLOADS[FLY]["res"]
    rows = W.path(FLY, 30)
    S_ = (rx // 16) * (ry // 16)
    o = O.Frame(rx, ry, words, root, depth, nthreads=16)
    try:
        with S.Context(0) as c:
            c.upload_octree(words, root, depth)
            c.resize(rx, ry)
            c.set_pretrace(3, 512, only_queued=0, predict=0)
            longest_main, longest_pre = [], []
            for f in range(30):
                p = W.frame_params(FLY, rows, f)
                print(f"pretrace flythrough frame {f}", flush=True)
                c.ray_stats(reset=True)
                c.pretrace_stats(reset=True)
                c.frame_warp(p)
                o.frame_warp(O.params_from(p))
                c.sync()
                assert_equal_u32(c.screen(0), o.screen[0], f"frame {f} plane 0")
                ids, oids = c.ids(), o.ids
                assert_equal_u32(ids[:2 * S_ + int(oids[0])], oids[:2 * S_ + int(oids[0])], f"frame {f} ids")
                assert_equal_u32(c.image(), o.image, f"frame {f} image")
                if f >= 5:
                    longest_main.append(c.ray_stats()[1])
                    longest_pre.append(c.pretrace_stats()[2])
            print("longest ray traced in the frame", longest_main, "ahead", longest_pre)
            assert np.median(longest_main) < np.median(longest_pre)
    finally:
        o.close()


def test_pretrace_timing_and_resize(scene1):
    """Per-frame timing (mode 2) covers the pretrace; a resize and an octree
    upload between frames start the sets over; stage timing (mode 1) runs
    frames without a pretrace.  Frames still equal a context without one."""
    words, root = scene1
    rx, ry = CONFIG1["res"]
    pr = path(CONFIG1, 12)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, CONFIG1["depth"])
            c.resize(rx, ry)
        b.set_pretrace(0)
        a.set_pretrace(only_queued=0)
        a.enable_timing(mode=2)
        for f in range(6):
            p = params(CONFIG1, pr, f)
            a.frame_warp(p)
            b.frame_warp(p)
            a.sync()
            assert a.frame_time() > 0.0
        _same(a, b, "timed frames")
        for c in (a, b):
            c.resize(rx - 64, ry - 32)
            c.upload_octree(words, root, CONFIG1["depth"])
        a.enable_timing(mode=1)
        for f in range(6, 12):
            p = params(CONFIG1, pr, f - 6)
            a.pretrace_stats(reset=True)
            a.frame_warp(p)
            b.frame_warp(p)
            a.sync()
            assert a.pretrace_stats()[0] == 0
        _same(a, b, "after resize, stage timing")
        a.enable_timing(False)
        for f in range(2, 8):
            p = params(CONFIG1, pr, f)
            a.frame_warp(p)
            b.frame_warp(p)
        a.sync()
        assert a.pretrace_stats()[1] > 0
        _same(a, b, "after resize, pretrace again")
    finally:
        a.close()
        b.close()


def test_pretrace_only_for_queued_frames():
    """With only_queued (the default), frames issued back to back are
    pretraced and a frame issued to an idle GPU is not; frame completion
    times (timing mode 3) cover the frames issued.  Frames equal a context
    without a pretrace (the headline workload, 1920x1024)."""
    words, root, depth, _ = W.octree(FLY)
    rx, ry = W.WORKLOADS[FLY]["res"]
    rows = W.path(FLY, 30)
    P = [W.frame_params(FLY, rows, f) for f in range(30)]
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, depth)
            c.resize(rx, ry)
        b.set_pretrace(0)
        assert a.get_pretrace().only_queued == 1
        for f in range(4):
            a.frame_warp(P[f])
            b.frame_warp(P[f])
        a.sync()
        a.pretrace_stats(reset=True)
        a.enable_timing(mode=3)
        for f in range(4, 20):  # back to back
            a.frame_warp(P[f])
        ts = a.frame_intervals()
        a.enable_timing(False)
        assert len(ts) == 15 and (ts > 0).all()
        pre_rays, merged, _ = a.pretrace_stats(reset=True)
        print("queued frames: rays traced ahead", pre_rays, "merged", merged, "intervals", ts.round(3), flush=True)
        assert merged > 0
        for f in range(4, 20):
            b.frame_warp(P[f])
        _same(a, b, "back-to-back frames")
        for f in range(20, 26):  # each frame waits for the one before
            a.frame_warp(P[f])
            a.sync()
            b.frame_warp(P[f])
        assert a.pretrace_stats()[0] == 0
        _same(a, b, "synchronised frames")
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("max_cnn_holes", [64, 0])
def test_pretrace_learned_frames_equal_plain_learned_frames(scene1, max_cnn_holes):
    """Learned frames (gate + CNN) with a pretrace equal learned frames
    without one: the gate's rays, the holes and the tile take the traced-ahead
    results; the CNN's input is plane 0 before any ray."""
    from svo_amd import learned as L
    words, root = scene1
    rx, ry = CONFIG1["res"]
    pr = path(CONFIG1, 14)
    cnn = S.Cnn(0)
    L.load(cnn, L.WEIGHTS)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, CONFIG1["depth"])
            c.resize(rx, ry)
        a.set_pretrace(3, 64, only_queued=0, predict=0)
        b.set_pretrace(0)
        for f in range(14):
            print(f"pretrace learned frame {f} max_cnn_holes {max_cnn_holes}", flush=True)
            p = params(CONFIG1, pr, f)
            a.pretrace_stats(reset=True)
            a.frame_warp_learned(cnn, p, max_cnn_holes)
            b.frame_warp_learned(cnn, p, max_cnn_holes)
            a.sync()
            if f >= 2:
                assert a.pretrace_stats()[1] > 0
            if f % 3 == 1 or f == 13:
                _same(a, b, f"learned frame {f}")
    finally:
        a.close()
        b.close()
        cnn.close()


def test_pretrace_rejects_bad_settings(scene1):
    with S.Context(0) as c:
        for args in [(1, 512), (9, 512), (-1, 512), (3, 0), (3, (1 << 22) + 1), (3, 512, 33), (3, 512, -1),
                     (3, 512, 0, -1), (3, 512, 0, 0, 3), (3, 512, 0, 0, 0), (3, 512, 0, 0, 128),
                     (3, 512, 0, 0, 8, 1, -1), (3, 512, 0, 0, 8, 1, 4), (3, 512, 0, 0, 8, 1, 2, -1),
                     (3, 512, 0, 0, 8, 1, 2, 0, -1), (3, 512, 0, 0, 8, 1, 2, 0, 5)]:
            with pytest.raises(S.SvoError):
                c.set_pretrace(*args)


@pytest.mark.parametrize("predict,long_trips,graph,sets,whole,qskip", [
    (1, 1, True, 3, 0, 0), (2, 1, False, 4, 0, 2), (2, 64, True, 6, 0, 4), (3, 1, True, 8, 0, 3),
    (3, 128, False, 2, 0, 0), (2, 1, True, 4, 256, 1), (3, 1, False, 6, 64, 0), (2, 1, True, 2, 1, 4)])
def test_pretrace_predicted_holes_equal_plain_frames(scene1, predict, long_trips, graph, sets, whole, qskip):
    """Predicted-hole pretrace (frame g's hole quads from frame g-k's final
    plane 0 through the cameras in between; with whole > 0 the blocks of long
    known rays traced whole by a second launch): frames equal a context
    without a pretrace, every traced or merged ray is one of the frame's, and
    from frame k + 2 on the pretrace carries the hole rays -- the frame itself
    traces few of them."""
    words, root = scene1
    rx, ry = CONFIG1["res"]
    n = 24
    pr = path(CONFIG1, n)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, CONFIG1["depth"])
            c.resize(rx, ry)
            c.set_graph(graph)
        a.set_pretrace(sets, long_trips, only_queued=0, predict=predict, whole_trips=whole, queue_skip=qskip)
        got = a.get_pretrace()
        assert (got.predict, got.whole_trips, got.queue_skip) == (predict, whole, qskip)
        b.set_pretrace(0)
        in_frame, holes = 0, 0
        for f in range(n):
            print(f"predicted pretrace k={predict} long={long_trips} graph={graph} sets={sets} whole={whole} "
                  f"qskip={qskip} frame {f}", flush=True)
            if f == n // 2:  # the sets move to other streams mid-path
                a.set_pretrace(queue_skip=(qskip + 1) % 5)
            p = params(CONFIG1, pr, f)
            a.ray_stats(reset=True)
            a.pretrace_stats(reset=True)
            a.frame_warp(p)
            b.frame_warp(p)
            a.sync()
            rays, _ = a.ray_stats()
            pre_rays, merged, _ = a.pretrace_stats()
            H = int(b.ids()[0])
            _, _, ex, ey = S.tile_schedule(rx, ry, f)
            assert rays + merged == H + ex * ey, f"frame {f}"
            if f >= predict + 2:
                in_frame += rays
                holes += H
            if f % 4 == 3 or f == n - 1:
                _same(a, b, f"frame {f}")
        print("hole rays traced in the frame", in_frame, "of", holes, flush=True)
        if long_trips == 1:  # the prediction misses a few holes, more the further back it starts
            assert in_frame <= (0.1 if predict <= 2 else 0.2) * holes
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("sets,whole,split,lanes", [(4, 1024, 0, 8), (3, 0, 2048, 1)])
def test_pretrace_predicted_holes_flythrough_vs_oracle(sets, whole, split, lanes):
    """The headline workload with predicted holes (k = 2; with whole > 0
    the blocks of rays of that many trips and more whole; with split > 0
    the rays of at least that many known trips in waves of their own,
    `lanes` per wave -- the bench's tuned setting): frames 0..29 against the
    oracle (plane 0, ids, image); the frame's own launch traces under 5% of
    the hole rays from frame 4 on.  A jump in the frame counter (no snapshot
    of frame g-2) falls back to the long blocks and still equals the
    oracle."""
    words, root, depth, _ = W.octree(FLY)
    rx, ry = W.WORKLOADS[FLY]["res"]
    rows = W.path(FLY, 30)
    S_ = (rx // 16) * (ry // 16)
    o = O.Frame(rx, ry, words, root, depth, nthreads=16)
    try:
        with S.Context(0) as c:
            c.upload_octree(words, root, depth)
            c.resize(rx, ry)
            c.set_pretrace(sets, 1, split_trips=split, split_lanes=lanes, only_queued=0, predict=2,
                           whole_trips=whole)
            in_frame, holes = 0, 0
            for f in range(30):
                fi = f if f < 20 else f + 100  # frame counters 120.. after frame 19
                p = W.frame_params(FLY, rows, f, frame_index=fi)
                print(f"predicted pretrace flythrough sets={sets} whole={whole} split={split}/{lanes} frame {f} "
                      f"(counter {fi})", flush=True)
                c.ray_stats(reset=True)
                c.pretrace_stats(reset=True)
                c.frame_warp(p)
                o.frame_warp(O.params_from(p))
                c.sync()
                assert_equal_u32(c.screen(0), o.screen[0], f"frame {f} plane 0")
                ids, oids = c.ids(), o.ids
                assert_equal_u32(ids[:2 * S_ + int(oids[0])], oids[:2 * S_ + int(oids[0])], f"frame {f} ids")
                assert_equal_u32(c.image(), o.image, f"frame {f} image")
                if 4 <= f < 20:
                    in_frame += c.ray_stats()[0]
                    holes += int(oids[0])
            print("hole rays traced in the frame", in_frame, "of", holes, flush=True)
            assert in_frame < 0.05 * holes
    finally:
        o.close()


@pytest.mark.parametrize("max_cnn_holes,gate_depth", [(64, 0.0), (0, 0.0), (64, 40.0)])
def test_pretrace_predicted_gate_learned_frames(scene1, max_cnn_holes, gate_depth):
    """Learned frames whose pretrace also predicts the gate (tiles with more
    than max_cnn_holes residual holes, and with the depth rule the tiles
    holding far reprojected pixels) equal learned frames without a pretrace;
    the pretrace carries most of the gate's rays."""
    from svo_amd import learned as L
    words, root = scene1
    rx, ry = CONFIG1["res"]
    pr = path(CONFIG1, 16)
    cnn = S.Cnn(0)
    L.load(cnn, L.WEIGHTS)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, CONFIG1["depth"])
            c.resize(rx, ry)
            c.set_gate_depth(gate_depth)
        a.set_pretrace(4, 1, only_queued=0, predict=2)
        b.set_pretrace(0)
        in_frame, total = 0, 0
        for f in range(16):
            print(f"predicted gate, learned frame {f} max_cnn_holes {max_cnn_holes} depth {gate_depth}", flush=True)
            p = params(CONFIG1, pr, f)
            a.ray_stats(reset=True)
            a.pretrace_stats(reset=True)
            a.frame_warp_learned(cnn, p, max_cnn_holes)
            b.frame_warp_learned(cnn, p, max_cnn_holes)
            a.sync()
            rays, _ = a.ray_stats()
            _, merged, _ = a.pretrace_stats()
            if f >= 4:
                in_frame += rays
                total += rays + merged
            if f % 3 == 1 or f == 15:
                _same(a, b, f"learned frame {f}")
        print("rays traced in the frame", in_frame, "of", total, flush=True)
        assert in_frame < 0.5 * total
    finally:
        a.close()
        b.close()
        cnn.close()


def test_pretrace_predicted_holes_learned_frames(scene1):
    """Learned frames with predicted holes equal learned frames without a
    pretrace."""
    from svo_amd import learned as L
    words, root = scene1
    rx, ry = CONFIG1["res"]
    pr = path(CONFIG1, 14)
    cnn = S.Cnn(0)
    L.load(cnn, L.WEIGHTS)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, CONFIG1["depth"])
            c.resize(rx, ry)
        a.set_pretrace(3, 1, only_queued=0, predict=2)
        b.set_pretrace(0)
        for f in range(14):
            print(f"predicted pretrace learned frame {f}", flush=True)
            p = params(CONFIG1, pr, f)
            a.frame_warp_learned(cnn, p, 64)
            b.frame_warp_learned(cnn, p, 64)
            a.sync()
            if f % 3 == 1 or f == 13:
                _same(a, b, f"learned frame {f}")
    finally:
        a.close()
        b.close()
        cnn.close()
