"""The 64k-particle parity run against the f64 oracle, at full length.

Workload: the 64 x 32 x 32 dam break (65 536 fluid + 26 802 Akinci wall
particles, dx 5 mm, h = 2 dx, jitter +-0.01 dx from seed 0), 200 steps of
2.5e-5 s, on the device and in the oracle side by side.

Bars:
  - density RMS relative error of the fluid <= 1e-5 at every 10th step and at
    the end;
  - at steps 0, 100 and 200, on the device's own positions: the reference grid
    (keys, order, occupied cells, ranges) bit-exact, the unfiltered candidate
    lists of sampled particles bit-exact in the reference's order, and the
    device's r < 2h lists equal to the oracle's sets (pairs within 1e-6 h of
    the support edge excluded);
  - mass conserved exactly (FP64 device reduction).
"""
import os

import numpy as np
import pytest

import pyoracle as O
from pef import scenes
from helpers import gpu_world, oracle_scene, rel_rms

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

CFG = scenes.CHECK_64K


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _check_lists(W, prm, a, x, rng):
    G = O.World(prm, x, a["kind"], psi=a["psi"], mode=O.FAST, threads=_threads())
    G.rebuild_grid()
    g, go = W.grid(), G.grid()
    for k in ("keys", "order", "occupied", "ranges"):
        assert np.array_equal(g[k].astype(np.int64), go[k].astype(np.int64)), k
    h = prm["h"]
    for i in rng.choice(len(x), size=64, replace=False):
        cand = W.neighbours(int(i))
        assert np.array_equal(cand.astype(np.int64), G.neighbours(int(i)).astype(np.int64))
        if a["kind"][i] != 0:
            continue
        r = np.linalg.norm(x[cand] - x[i], axis=1)
        want = set(int(j) for j, rr in zip(cand, r) if rr < 2 * h and j != i and abs(rr - 2 * h) > 1e-6 * h)
        got = set(int(j) for j in W.neighbour_list(int(i)))
        edge = set(int(j) for j, rr in zip(cand, r) if abs(rr - 2 * h) <= 1e-6 * h)
        assert want - edge == got - edge


def test_check_64k_200_steps_against_the_oracle():
    W, parts = gpu_world(CFG)
    prm, a = oracle_scene(CFG)
    Wo = O.World(prm, a["x"], a["kind"], psi=a["psi"], mode=O.FAST, threads=_threads())
    assert np.array_equal(parts.x.view(np.uint64), a["x"].view(np.uint64))
    fl = a["kind"] == 0
    rng = np.random.default_rng(7)
    m0 = W.diagnostics()["total_mass"]
    worst = 0.0
    _check_lists(W, prm, a, W.positions(), rng)
    for k in range(1, CFG.steps + 1):
        W.step(CFG.dt)
        Wo.step(CFG.dt)
        if k % 10 == 0:
            e = rel_rms(W.densities()[fl], Wo.state()["rho"][fl])
            worst = max(worst, e)
            print(f"step {k}: density rms rel {e:.3e}", flush=True)
            assert e <= 1e-5, (k, e)
        if k in (100, CFG.steps):
            _check_lists(W, prm, a, W.positions(), rng)
    assert W.diagnostics()["total_mass"] == m0
    print(f"worst density rms rel over 200 steps: {worst:.3e}")
