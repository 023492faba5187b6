"""Dam-break quantities of interest on the device (validation.rs:63-117).

Bars:
  - column heights at rest equal the closed form (ny particles of area
    m / rho0 in each column's bin: height ny m / (rho0 bw), relative 1e-14)
    and the front equals the last column's x;
  - on the device's own positions after stepping, heights and bin centres
    are bit-identical to a per-particle restatement of validation.rs:66-92
    (written here as a plain loop), for several bin widths and ranges that
    cut through the fluid, and per scene of a two-scene batch;
  - wcsph_dam_break_front_at (validation.rs:415-439) returns the front of the
    world it stepped;
  - its 2-D inviscid dam break (dx 5 mm, h0 0.1 m, 400 fluid particles), run
    on the device and in the f64 oracle side by side: after 300 steps every
    fluid particle within 1e-3 dx, the same front and mean |dh| <= 0.01 h0.
    The flow is chaotic: in the oracle alone, positions scaled by
    1 + U(-1e-7, 1e-7) (FP32-sized noise) separate by 5-8 dx after 1000 steps
    and by 25-30 dx after 3000, with fronts 0-5 dx apart and mean |dh| of
    0.15-0.20 h0 over wet bins. At 1000, 2000 and 3000 steps the device is
    held to that spread: fronts within 6 dx, mean |dh| <= 0.3 h0.
"""
import math

import numpy as np
import pytest

import pef
import pyoracle as O
from pef import validation as V

pytestmark = pytest.mark.gpu

DX, H0, LRES, CH = 0.01, 0.1, 0.1, 0.5


def _heights_restated(x, kind, mass, rho0, x_min, x_max, bin_w):
    """validation.rs:66-92 particle by particle."""
    bw = max(bin_w, 1e-12)
    n_bins = max(int(max(math.ceil((x_max - x_min) / bw), 0)), 1)
    acc = [0.0] * n_bins
    area = mass / max(rho0, 1e-12)
    for xi, k in zip(x[:, 0].tolist(), kind.tolist()):
        if k != pef.FLUID or xi < x_min or xi >= x_max:
            continue
        b = int(math.floor((xi - x_min) / bw))
        if b < n_bins:
            acc[b] += area
    xc = np.array([x_min + (b + 0.5) * bw for b in range(n_bins)])
    return xc, np.array([a / bw for a in acc])


def _world(dx=DX, h0=H0, l_res=LRES, ch=CH):
    parts, dam_x = V.seed_dam_break_particles(dx, h0, l_res, ch)
    return pef.WcsphWorld(parts, V.wcsph_dam_break_params()), parts, dam_x


def test_column_heights_closed_form_at_rest():
    W, parts, dam_x = _world()
    nx, ny = round(LRES / DX), round(H0 / DX)
    xc, h = W.column_heights(0.0, CH, DX)
    assert len(xc) == math.ceil(CH / DX)
    area = parts.mass / parts.rest_density
    want = ny * area / DX
    assert np.all(np.abs(h[:nx] - want) <= 1e-14 * want)
    assert np.all(h[nx:] == 0.0)
    assert np.array_equal(xc, np.array([0.0 + (b + 0.5) * DX for b in range(len(xc))]))
    # the front sits in the bin of the last column: centre = that column's x
    xf = W.front_x_height(DX, H0, CH, 0.02)
    assert xf == pytest.approx(dam_x - 0.5 * DX, abs=1e-12)
    assert V.fluid_front_x_height(W, DX, H0, CH) == xf
    # a threshold above the column height leaves everything dry
    assert W.front_x_height(DX, H0, CH, 1.5) == -math.inf
    W.close()


@pytest.mark.parametrize("bw_over_dx,x_min,x_max", [(1.0, -2 * DX, CH + 2 * DX), (0.37, 0.013, 0.21),
                                                     (3.0, -0.05, 0.33), (1.0, 0.05, 0.051)])
def test_column_heights_match_restatement_after_steps(bw_over_dx, x_min, x_max):
    W, parts, _ = _world()
    W.step(2e-5, 400)
    x = W.positions()
    bw = bw_over_dx * DX
    xc, h = W.column_heights(x_min, x_max, bw)
    xr, hr = _heights_restated(x, parts.kind, parts.mass, parts.rest_density, x_min, x_max, bw)
    assert np.array_equal(xc.view(np.uint64), xr.view(np.uint64))
    assert np.array_equal(h.view(np.uint64), hr.view(np.uint64))
    W.close()


def test_column_heights_per_scene_of_a_batch():
    p0, _ = V.seed_dam_break_particles(DX, H0, LRES, CH)
    p1, _ = V.seed_dam_break_particles(DX, 0.8 * H0, 0.6 * LRES, 0.7 * CH)
    prm = V.wcsph_dam_break_params()
    W = pef.WcsphWorld([p0, p1], [prm, prm])
    W.step(2e-5, 200)
    x = W.positions()
    for k, p in enumerate((p0, p1)):
        a, n = W.scene_range(k)
        assert n == len(p)
        xr, hr = _heights_restated(x[a:a + n], p.kind, p.mass, p.rest_density, -2 * DX, CH, DX)
        xc, h = W.column_heights(-2 * DX, CH, DX, scene=k)
        assert np.array_equal(h.view(np.uint64), hr.view(np.uint64)), k
        assert np.array_equal(xc.view(np.uint64), xr.view(np.uint64)), k
    W.close()


def test_column_heights_argument_errors():
    W, _, _ = _world()
    with pytest.raises(pef.PefError):
        W.column_heights(0.0, CH, DX, scene=1)
    with pytest.raises(pef.PefError):
        W.column_heights(float("nan"), CH, DX)
    # an empty or reversed range still has one bin (n_bins >= 1), as in the reference
    xc, h = W.column_heights(0.3, 0.2, DX)
    assert len(xc) == 1 and h[0] == 0.0
    W.close()


def test_wcsph_dam_break_front_at_is_the_stepped_world():
    t_end = 0.01
    xf, rel, t = V.wcsph_dam_break_front_at(DX, H0, LRES, t_end, CH)
    dt = min(0.2 * DX / 50.0, 2e-4)
    W, _, dam_x = _world()
    W.step(dt, max(math.ceil(t_end / dt), 1))
    assert t == W.t and xf == W.front_x_height(DX, H0, CH, 0.02)
    assert rel == V.ritter_front_rel_error(xf, dam_x, 9.81, H0, t)
    W.close()


def test_dam_break_front_and_heights_match_oracle():
    dx, h0, l_res, ch = 0.005, 0.1, 0.1, 0.4
    dt = min(0.2 * dx / 50.0, 2e-4)
    parts, dam_x = V.seed_dam_break_particles(dx, h0, l_res, ch)
    W = pef.WcsphWorld(parts, V.wcsph_dam_break_params())
    prm = dict(dim=2, rest_density=1000.0, mass=parts.mass, h=parts.h, sound_speed=50.0, visc_alpha=0.0,
               xsph_eps=0.0, gravity=(0.0, -9.81, 0.0), adapt_sound_speed=False)
    Wo = O.World(prm, parts.x, parts.kind, psi=parts.psi, threads=16)
    fl = parts.kind == pef.FLUID
    area = parts.mass / parts.rest_density
    lo, hi = -2 * dx, ch + 2 * dx
    done = 0
    for target in (300, 1000, 2000, 3000):
        W.step(dt, target - done)
        Wo.step(dt, target - done)
        done = target
        assert W.t == pytest.approx(Wo.t, rel=1e-12)
        xo_p = Wo.state()["x"]
        xo, ho = _heights_restated(xo_p, parts.kind, parts.mass, parts.rest_density, lo, hi, dx)
        xf_o = V.front_x_from_height_field(list(zip(xo, ho)), h0, 0.02)
        xf = W.front_x_height(dx, h0, ch, 0.02)
        xc, h = W.column_heights(lo, hi, dx)
        wet = (h >= 0.02 * h0) | (ho >= 0.02 * h0)
        dh = float(np.mean(np.abs(h[wet] - ho[wet]))) / h0
        drift = float(np.abs(W.positions() - xo_p)[fl].max()) / dx
        print(f"step {target} t {W.t:.3f}: front device {xf:.4f} oracle {xf_o:.4f} "
              f"(ritter {dam_x + V.ritter_front(9.81, h0, W.t):.4f}); mean |dh|/h0 {dh:.3f}; "
              f"max |dx|/dx {drift:.2e}", flush=True)
        assert np.array_equal(xc, xo)
        # every fluid particle is binned once: sum(h) bw = n_fluid area
        assert float(np.sum(h)) * dx == pytest.approx(fl.sum() * area, rel=1e-12)
        if target == 300:
            assert drift <= 1e-3
            assert xf == xf_o and dh <= 0.01
        else:
            assert abs(xf - xf_o) <= 6 * dx + 1e-12
            assert dh <= 0.3
    W.close()
