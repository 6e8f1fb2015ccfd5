"""Train the learned hole filler on frame pairs generated along seeded
flythroughs of both workloads and save the weights.

    python -m svo_amd.train_holefill [--pairs 20000] [--out PATH]

Pairs alternate between the 512x256 depth-8 flythrough and the 1920x1024
depth-11 flythrough, each path replayed under successive seed offsets."""
import argparse
import sys
import time

from . import Cnn
from . import learned as L


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--crops", type=int, default=16)
    ap.add_argument("--crop", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--train-max-holes", type=int, default=256,
                    help="gate threshold of the training frames (more CNN tiles -> more loss pixels)")
    ap.add_argument("--out", default=L.WEIGHTS)
    a = ap.parse_args()
    plan = []
    left, k = a.pairs, 0
    while left > 0:
        name, n = ("flythrough_256", 64) if k % 2 == 0 else ("flythrough_2048", 300)
        n = min(n, left)
        plan.append((name, k // 2, n))
        left -= n
        k += 1
    cnn = Cnn(0, seed=a.seed)
    t = time.time()
    losses = L.train(cnn, plan, a.crops, a.crop, a.lr, a.lr / 10, log=lambda m: print(m, flush=True),
                     log_every=500, max_cnn_holes=a.train_max_holes)
    dt = time.time() - t
    got = [x for x in losses if x is not None]
    L.save(cnn, a.out, pairs=a.pairs, crops=a.crops, crop=a.crop, seconds=dt, final_loss=got[-1],
           train_max_holes=a.train_max_holes)
    print(f"trained {a.pairs} pairs in {dt:.1f}s, final loss {got[-1]:.5f} -> {a.out}", flush=True)
    for name in ("flythrough_256", "flythrough_2048"):
        allp, holes = L.evaluate(cnn, name, n_frames=40)
        print(f"held-out {name}: PSNR all {allp}  holes {holes}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
