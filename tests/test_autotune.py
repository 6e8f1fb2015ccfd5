"""Autotuner search logic on the CPU (synthetic cost functions) and, on the
GPU, a tuned context still matching the oracle bit-exactly."""
import pytest

import oracle as O
import svo_amd as S
from svo_amd import autotune as A
from helpers import CONFIG1, assert_equal_u32, params, path, scene_octree


def test_grid_search_finds_minimum():
    space = {"a": [1, 2, 3], "b": [10, 20], "c": [0, 5]}
    cost = lambda cfg: (cfg["a"] - 2) ** 2 + abs(cfg["b"] - 20) + 0.1 * cfg["c"]
    seen = []
    best, ms = A._search(space, False, lambda cfg: seen.append(cfg) or cost(cfg))
    assert best == {"a": 2, "b": 20, "c": 0} and ms == 0.0
    assert len(seen) == 12


def test_coordinate_descent_on_separable_cost():
    space = {"a": [4, 1, 2, 3], "b": [7, 5, 6], "c": [9, 8]}
    cost = lambda cfg: abs(cfg["a"] - 3) + abs(cfg["b"] - 7) + abs(cfg["c"] - 9)
    best, ms = A._search(space, True, cost)
    assert best == {"a": 3, "b": 7, "c": 9} and ms == 0
    # the start point is each knob's second value
    first = []
    A._search(space, True, lambda cfg: first.append(dict(cfg)) or 0.0)
    assert first[0] == {"a": 1, "b": 5, "c": 8}


def test_full_space_adds_long_first():
    assert "long_trips" in A.FULL_SPACE and "long_trips" not in A.DEFAULT_SPACE
    assert all(k in A.FULL_SPACE for k in A.DEFAULT_SPACE)


def test_warp_space_adds_pretrace():
    assert "pre_sets" in A.WARP_SPACE and "pre_long" in A.WARP_SPACE and "pre_sets" not in A.DEFAULT_SPACE
    assert 0 in A.WARP_SPACE["pre_pred"] and max(A.WARP_SPACE["pre_pred"]) <= 3
    assert all(0 <= q <= 4 for q in A.WARP_SPACE["pre_qskip"]) and A.WARP_SPACE["pre_qskip"][1] == 3
    assert all(k in A.WARP_SPACE for k in A.DEFAULT_SPACE)


class _Recorder:
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        return lambda *a, **k: self.calls.append((name, a, k))


def test_apply_a_partial_setting():
    """A setting of some knobs only (the divergence stress tunes the
    pretrace alone) keeps the ray-kernel knobs it does not name."""
    r = _Recorder()
    A.apply(r, {"pre_sets": 3, "pre_qskip": 2})
    assert ("set_tuning", (-1, -1, -1, -1), {}) in r.calls
    pre = [c for c in r.calls if c[0] == "set_pretrace"]
    assert len(pre) == 1 and pre[0][2]["sets"] == 3 and pre[0][2]["queue_skip"] == 2 and pre[0][2]["predict"] is None
    assert pre[0][2]["wg_per_cu"] is None


def test_learned_space_adds_cnn_workgroups():
    assert "cnn_wgs" in A.LEARNED_SPACE and "cnn_wgs" not in A.DEFAULT_SPACE
    assert all(k in A.LEARNED_SPACE for k in A.DEFAULT_SPACE)
    assert A.LEARNED_SPACE["cnn_wgs"][1] == 128  # the descent starts at half the CUs


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["warp", "full"])
def test_tuned_context_stays_bitexact(mode):
    words, root = scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])
    rx, ry = CONFIG1["res"]
    pr = path(CONFIG1, 10)
    ps = [params(CONFIG1, pr, f) for f in range(10)]
    space = {"spread": [1, 64], "refill": [8, 32], "wg_per_cu": [4], "iters": [16, 64], "lds_cache_words": [0, 736]}
    if mode == "full":
        space["long_trips"] = [0, 16]
    else:
        space["pre_sets"] = [6, 3]
        space["pre_pred"] = [0, 2]
        space["pre_long"] = [256, 1]
        space["pre_qskip"] = [0, 3]
    c = S.Context(0)
    o = O.Frame(rx, ry, words, root, 8, nthreads=8)
    try:
        c.upload_octree(words, root, 8)
        c.resize(rx, ry)
        best, ms, trials = A.tune(c, ps, warmup=4, mode=mode, space=space)
        print("tuned", mode, best, ms, flush=True)
        assert all(best[k] in space[k] for k in space) and ms > 0 and len(trials) >= 4
        # the tuned context replays the path from a clear frame
        for f in range(6):
            print(f"tuned {mode} frame {f}", flush=True)
            (c.frame_warp if mode == "warp" else c.frame_full)(ps[f])
            c.sync()
            (o.frame_warp if mode == "warp" else o.frame_full)(O.params_from(ps[f]))
            if mode == "warp":
                assert_equal_u32(c.screen(), o.screen, f"frame {f} screen")
            else:
                assert_equal_u32(c.screen(0), o.screen[0], f"frame {f} screen")
            assert_equal_u32(c.image(), o.image, f"frame {f} image")
    finally:
        c.close()
        o.close()


@pytest.mark.gpu
def test_tuned_learned_frames_equal_default_knobs():
    """Learned-frame tuning (gate + CNN frames) picks knobs that change only
    the schedule: the tuned context's frames equal a default context's."""
    from svo_amd import learned as L
    words, root = scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])
    rx, ry = CONFIG1["res"]
    pr = path(CONFIG1, 10)
    ps = [params(CONFIG1, pr, f) for f in range(10)]
    space = {"refill": [8, 32], "wg_per_cu": [2, 4], "iters": [16, 64], "lds_cache_words": [0, 736]}
    cnn = S.Cnn(0)
    L.load(cnn, L.WEIGHTS)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, 8)
            c.resize(rx, ry)
        best, ms, trials = A.tune(a, ps, warmup=4, mode="learned", space=space, cnn=cnn, max_cnn_holes=64)
        print("tuned learned", best, ms, flush=True)
        assert all(best[k] in space[k] for k in space) and ms > 0
        for f in range(6):
            a.frame_warp_learned(cnn, ps[f], 64)
            b.frame_warp_learned(cnn, ps[f], 64)
            a.sync()
            b.sync()
            assert_equal_u32(a.screen(), b.screen(), f"frame {f} screen")
            assert_equal_u32(a.image(), b.image(), f"frame {f} image")
    finally:
        a.close()
        b.close()
        cnn.close()
