"""Train the capacity model on the GPU and write yggdrax_amd/data/capacity_mlp.npz.

    python -m yggdrax_amd.train_capacity [--out PATH] [--steps N]

Training data: 200 seeded systems (seeds 1000-1199) from the uniform-cube,
Plummer and clustered generators at sizes up to 2^16, 2^20 and 2^24
points; calibration on 30 held-out systems (seeds 2000-2029). Every target
comes from the project's own walk.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch


def main():
    from yggdrax_amd import learned

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=learned.DEFAULT_WEIGHTS)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    args = ap.parse_args()
    t0 = time.time()
    x, y, m, _ = learned.collect(learned.training_systems(range(1000, 1200)))
    print(f"[capacity] {x.shape[0]} node samples from 200 systems ({time.time() - t0:.1f} s)",
          flush=True)
    
    params, scale = learned.train(x, y, m, steps=args.steps, batch=args.batch, lr=args.lr,
                                  log=lambda s: print(s, flush=True))
    model = learned.CapacityModel(params, scale)
    _, _, _, held = learned.collect(learned.training_systems(range(2000, 2030)), keep_trees=True)
    learned.calibrate(model, held)
    fit, over = 0, []
    for r in held:
        hint = model.capacity_hint(r["tree"], r["geom"])
        t = r["totals"]
        fit += int(hint["far"] >= t[0] and hint["near"] >= t[1] and hint["queue"] >= t[2])
        over.append([hint["far"] / max(t[0], 1), hint["near"] / max(t[1], 1)])
    model.meta = dict(train_seeds="1000-1199", calib_seeds="2000-2029", steps=args.steps,
                      held_out_fit=f"{fit}/{len(held)}", created=time.strftime("%Y-%m-%d"))
    model.save(args.out)
    print(json.dumps(dict(scale=model.scale.tolist(), calib=model.calib.tolist(),
                          margin=model.margin.tolist(),
                          median_overalloc=np.median(np.array(over), axis=0).round(3).tolist(),
                          held_out_fit=f"{fit}/{len(held)}", seconds=round(time.time() - t0, 1),
                          out=args.out)), flush=True)


if __name__ == "__main__":
    sys.exit(main())
