"""Benchmark of the MI355X WCSPH step.

Workload: the 1,048,576-particle dam break (128 x 64 x 128 fluid lattice at
dx = 0.005 m, h = 2 dx, 163,842 Akinci wall particles, tank 1.6 x 0.32 x 0.8 m,
dt = 2.5e-5 s), FP32 arithmetic with FP64 position accumulation. One "step" is
one WcsphWorld::step of the whole scene, device-resident.

Prints one JSON line: particle-updates/s over the timed steps, the roofline of
the dominant kernel (algorithmic HBM bytes per launch / its HIP-event duration
against 8 TB/s), and the CPU f64 oracle timed on a bounded sample of the same
workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("physics-engine_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-updates/s (1M-particle SPH) + surrogate speedup at ≤0.5% density err"


def kernel_bytes(n_fluid, n_bnd, list_entries_total):
    """Algorithmic HBM bytes per launch of the step kernels.

    Own-particle streams only (the per-particle reads and writes each kernel
    makes once) plus the neighbour-list entries it reads (4 B each); gathers of
    neighbour state are served from cache and are not counted.
    """
    E = list_entries_total
    return {
        # fluid: META A2 B2 CNT ID CELL DISP X Y Z read (76) + list; IDF ACC BF X Y Z A B DISP written (96)
        # boundary: META A2 B2 CNT ID read (44) + list; IDF ACC BF written (36)
        "force": 172 * n_fluid + 80 * n_bnd + 4 * E,
        # fluid: A META CNT read + list, A2 written; boundary: A read, A2 written
        "density": 40 * n_fluid + 32 * n_bnd + 4 * E,
        # fluid: A2 B B2 META CNT read + list, B2 written; boundary: A2 B B2 read, B2 written
        "xsph": 72 * n_fluid + 64 * n_bnd + 4 * E,
    }


def cpu_baseline(cfg, seconds_hint=10.0):
    """f64 oracle, one thread, on the same scene for one step (bounded sample)."""
    import pyoracle as O
    from helpers import oracle_scene

    prm, a = oracle_scene(cfg)
    W = O.World(prm, a["x"], a["kind"], psi=a["psi"], mode=O.FAST, threads=1)
    W.step(cfg.dt, 1)  # the first step pays no extra setup; time the second
    t0 = time.perf_counter()
    W.step(cfg.dt, 1)
    dt = time.perf_counter() - t0
    return {"value": cfg.n_fluid / dt, "unit": "particle-updates/s", "cores": 1, "kind": "port",
            "sample": f"{cfg.name}: 1 step of the same scene, f64 oracle (reference op order, zero-contribution "
                      f"candidates skipped, bit-identical to the reference-structured path), 1 thread, "
                      f"{dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="dam_break_1m")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--skin", type=float, default=0.0)
    ap.add_argument("--lpp", type=int, default=0, help="lanes per particle of the pair kernels (0 = default)")
    ap.add_argument("--block", type=int, default=0, help="workgroup size of the pair kernels (0 = default)")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="also run one plain copy kernel of this size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU sharding is a design "
              f"note, not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pef
    from pef import scenes

    cfg = scenes.CONFIGS[args.config]
    t_setup = time.perf_counter()
    parts, params = scenes.build(cfg)
    W = pef.WcsphWorld(parts, params, pef.Options(skin=args.skin, lanes_per_particle=args.lpp, block_size=args.block))
    n_fluid = parts.fluid_count()
    n_bnd = len(parts) - n_fluid
    t_setup = time.perf_counter() - t_setup
    print(f"# setup {t_setup:.1f} s: {n_fluid} fluid + {n_bnd} boundary particles", file=sys.stderr, flush=True)

    W.step(cfg.dt, args.warmup)
    t0 = time.perf_counter()
    W.step(cfg.dt, args.steps)  # synchronous: returns after the device finished
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = n_fluid * args.steps / elapsed
    print(f"# timed {args.steps} steps: {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)

    mean_entries, rebuilds0 = W.list_stats()
    diag = W.diagnostics()

    # per-kernel HIP-event durations on the world's stream, same workload
    W.set_profiling(True)
    W.step(cfg.dt, args.profile_steps)
    times = W.kernel_times()
    W.set_profiling(False)
    _, rebuilds1 = W.list_stats()
    per_launch_ms = {k: v / args.profile_steps for k, v in times.items()}
    E = mean_entries * (n_fluid + n_bnd)
    kb = kernel_bytes(n_fluid, n_bnd, E)
    dom = max(("force", "density", "xsph"), key=lambda k: per_launch_ms.get(k, 0.0))
    dom_ms = per_launch_ms[dom]
    achieved = kb[dom] / (dom_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": f"k_{dom}",
            "bytes_per_launch": int(kb[dom]), "avg_launch_ms": round(dom_ms, 5)}
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name}: {n_fluid} fluid + {n_bnd} Akinci wall particles, dx {cfg.dx}, "
                               f"h {cfg.h}, dt {cfg.dt}, tank {cfg.tank}, jitter 0.01 dx seed {cfg.seed}",
                   "n_fluid": n_fluid, "n_boundary": n_bnd, "skin": args.skin or 0.08,
                   "mean_list_entries": round(mean_entries, 1), "list_rebuilds_in_profile": rebuilds1 - rebuilds0,
                   "surrogate_speedup": None},
        "roofline": roof,
        "kernel_ms": {k: round(v, 5) for k, v in per_launch_ms.items()},
        "tiles": W.tile_info(),
        "mean_density_error": diag["mean_density_error"],
    }
    if args.calibration_copy_mib:
        import ctypes as C
        pef.lib().pefx_calibration_copy.restype = C.c_int32
        pef.lib().pefx_calibration_copy.argtypes = [C.c_uint64]
        assert pef.lib().pefx_calibration_copy(args.calibration_copy_mib << 20) == 0
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
