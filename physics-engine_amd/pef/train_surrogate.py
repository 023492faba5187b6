"""Train the learned force surrogate on states of the plain step.

Recipe of the 1M dam break: 2 000 000 samples of fluid particles (features
and the plain force's a - g) drawn from runs of the 1 048 576-particle scene
at 20 jitter seeds, each sampled at geometrically spaced states over its
first `--horizon` steps; 2 000 Adam steps at batch 65 536 with the learning rate
stepped down over the run; a held-out seed measures the validation loss.
Everything runs on the device through libpef (pef_world_sample,
pef_dataset_*, pef_mlp_train).

    python physics-engine_amd/pef/train_surrogate.py [--out PATH] [--summary PATH]

writes the network (parameters and normalisation, .npz) and a short text
summary of the run.
"""
import argparse
import dataclasses
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import pef  # noqa: E402
from pef import scenes  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "models", "surrogate_dam_break_1m.npz")


def collect(cfg, seeds, per_seed, horizon, snapshots, ds, log):
    """Sample per_seed fluid particles of each seed's run at about `snapshots`
    states over [0, horizon] steps (step 0, then geometrically spaced)."""
    for s in seeds:
        t0 = time.perf_counter()
        c = dataclasses.replace(cfg, seed=s)
        parts, params = scenes.build(c)
        W = pef.WcsphWorld(parts, params)
        # geometric spacing: the acoustic relaxation of the jittered lattice in
        # the first tens of steps is sampled as densely as the slow collapse
        at = np.unique(np.concatenate([[0], np.geomspace(1, horizon, snapshots - 1).round().astype(int)]))
        counts = np.full(len(at), per_seed // len(at))
        counts[: per_seed - counts.sum()] += 1
        done = 0
        for k, (step, n) in enumerate(zip(at, counts)):
            if step > done:
                W.step(c.dt, int(step - done))
                done = int(step)
            W.sample(ds, int(n), seed=(s << 20) + k)
        W.close()
        log(f"seed {s}: {per_seed} samples over {horizon} steps in {time.perf_counter() - t0:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="dam_break_1m")
    ap.add_argument("--seeds", type=int, default=20, help="training seeds 0..seeds-1; seed `seeds` is held out")
    ap.add_argument("--samples", type=int, default=2_000_000)
    ap.add_argument("--val-samples", type=int, default=65536)
    ap.add_argument("--horizon", type=int, default=520, help="steps of each run the samples cover")
    ap.add_argument("--snapshots", type=int, default=16)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init-seed", type=int, default=2024)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--summary", default=None)
    args = ap.parse_args()
    lines = []

    def log(msg):
        print(msg, flush=True)
        lines.append(msg)

    cfg = scenes.CONFIGS[args.config]
    per_seed = args.samples // args.seeds
    t_all = time.perf_counter()
    ds = pef.Dataset(per_seed * args.seeds)
    collect(cfg, range(args.seeds), per_seed, args.horizon, args.snapshots, ds, log)
    val = pef.Dataset(args.val_samples)
    collect(cfg, [args.seeds], args.val_samples, args.horizon, args.snapshots, val, log)
    Xv, Tv = val.get()
    t_data = time.perf_counter() - t_all

    m = pef.Mlp(seed=args.init_seed)
    ds.fit_normalization(m)
    _, v0 = m.gradients(Xv, Tv)
    # learning rate lr for half the steps, lr / 3 for the next 30 %, lr / 10 for the rest
    plan = [(args.steps // 2, args.lr), (args.steps * 3 // 10, args.lr / 3)]
    plan.append((args.steps - sum(n for n, _ in plan), args.lr / 10))
    t0 = time.perf_counter()
    hist = np.concatenate([m.train(ds, n, batch=args.batch, lr=lr, seed=args.init_seed + 1) for n, lr in plan])
    t_train = time.perf_counter() - t0
    _, v1 = m.gradients(Xv, Tv)
    pred = m.predict(Xv)
    rel = float(np.sqrt(np.mean(np.sum((pred - Tv) ** 2, 1)) / np.mean(np.sum(Tv ** 2, 1))))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    m.save(args.out)
    log(f"data: {len(ds)} training + {len(val)} validation samples in {t_data:.1f} s")
    log(f"training: {args.steps} Adam steps at batch {args.batch}, lr {args.lr:g} / 3 / 10 at 50 % / 80 %, "
        f"{t_train:.1f} s ({1e3 * t_train / max(args.steps, 1):.2f} ms per step)")
    marks = [0, 9, 99, len(hist) // 2 - 1, len(hist) * 4 // 5 - 1, len(hist) - 1]
    log("batch loss (normalised MSE): " + ", ".join(f"step {k + 1}: {hist[k]:.4f}" for k in marks if k < len(hist)))
    log(f"validation (held-out seed {args.seeds}): loss {v0:.4f} -> {v1:.4f}; "
        f"RMS |a_pred - (a - g)| / RMS |a - g| = {rel:.4f}")
    log(f"saved {args.out}")
    if args.summary:
        with open(args.summary, "w") as f:
            f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
