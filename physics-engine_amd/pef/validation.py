"""Dam-break quantities of interest for the WCSPH world (validation.rs, oracles.rs).

The per-particle binning of `column_heights` runs on the device
(`WcsphWorld.column_heights`, C ABI `pef_world_column_heights`); everything
here works on the few hundred bin values it returns, or on scalars, in the
reference's order of operations:

  ritter_c0 / ritter_front / ritter_height ......... oracles.rs:37-63
  ritter_front_rel_error ........................... validation.rs:57-61
  front_x_from_height_field ........................ validation.rs:94-104
  ritter_height_profile_rel_error .................. validation.rs:119-140
  seed_dam_break_particles (2-D, Akinci bed/wall) .. validation.rs:142-190
  wcsph_dam_break_front_at ......................... validation.rs:415-439
"""
import math
from typing import Iterable, Optional, Sequence, Tuple

import numpy as np

import pef

DAM_BREAK_TAU_SAMPLES = (1.0, 2.0, 3.0, 4.0)  # validation.rs:25


def _rust_round(x: float) -> int:
    """`f64::round() as usize`: half away from zero, saturating at 0."""
    if not x > 0.0:
        return 0
    r = math.floor(x)
    if x - r >= 0.5:  # exact: x - floor(x) has no rounding error
        r += 1
    return int(r)


def ritter_c0(g: float, h0: float) -> float:
    """oracles.rs:37-39."""
    return math.sqrt(g * h0)


def ritter_front(g: float, h0: float, t: float) -> float:
    """oracles.rs:43-45: x_f = 2 c0 t."""
    return 2.0 * ritter_c0(g, h0) * t


def ritter_height(g: float, h0: float, x: float, t: float) -> float:
    """oracles.rs:48-63: Ritter's dry-bed rarefaction h(x, t), dam at x = 0."""
    if t <= 0.0:
        return h0 if x < 0.0 else 0.0
    c0 = ritter_c0(g, h0)
    x_l, x_f = -c0 * t, 2.0 * c0 * t
    if x <= x_l:
        return h0
    if x >= x_f:
        return 0.0
    c = (2.0 * c0 - x / t) / 3.0
    return (c * c) / g


def ritter_front_rel_error(x_sim: float, dam_x: float, g: float, h0: float, t: float) -> float:
    """validation.rs:57-61."""
    x_ref = dam_x + ritter_front(g, h0, t)
    return abs(x_sim - x_ref) / max(abs(x_ref), 1e-9)


def front_x_from_height_field(heights: Iterable[Tuple[float, float]], h0: float, h_frac: float = 0.02) -> float:
    """validation.rs:96-104: rightmost bin centre with h >= h_frac h0, -inf when dry."""
    thr = h_frac * h0
    xf = -math.inf
    for x, h in heights:
        if h >= thr:
            xf = x
    return xf


def ritter_height_profile_rel_error(heights: Iterable[Tuple[float, float]], dam_x: float, g: float, h0: float,
                                    t: float, h_frac: float = 0.02) -> Tuple[float, int]:
    """validation.rs:120-140: mean |h - h_ritter| / h0 over bins where either is wet."""
    thr = h_frac * h0
    s, n = 0.0, 0
    for x, h in heights:
        href = ritter_height(g, h0, x - dam_x, t)
        if h < thr and href < thr:
            continue
        s += abs(h - href) / max(h0, 1e-12)
        n += 1
    return (0.0 if n == 0 else s / n), n


def column_heights(world: "pef.WcsphWorld", x_min: float, x_max: float, bin_w: float,
                   scene: int = 0) -> Sequence[Tuple[float, float]]:
    """validation.rs:66-92 as (x_centre, height) pairs, binned on the device."""
    xc, h = world.column_heights(x_min, x_max, bin_w, scene)
    return list(zip(xc.tolist(), h.tolist()))


def fluid_front_x_height(world: "pef.WcsphWorld", dx: float, h0: float, channel_length: float,
                         h_frac: float = 0.02, scene: int = 0) -> float:
    """validation.rs:106-117."""
    return world.front_x_height(dx, h0, channel_length, h_frac, scene)


def seed_dam_break_particles(dx: float, h0: float, l_res: float, channel_length: float, h_over_dx: float = 2.2,
                             calibrate_cubic: bool = True, with_akinci_walls: bool = True):
    """validation.rs:142-190: 2-D column on the left, two-layer Akinci bed and
    left wall with psi = dx^2. Returns (Particles, dam_x)."""
    hsm = h_over_dx * dx
    rho0 = 1000.0
    if calibrate_cubic:
        mass = pef.mass_calibrated_to_rest(rho0, dx, hsm, pef.D2)
    else:
        mass = pef.mass_from_spacing(rho0, dx, pef.D2)
    parts = pef.Particles(pef.D2, rho0, mass, hsm)
    ny = max(_rust_round(h0 / dx), 6)
    nx = max(_rust_round(l_res / dx), 6)
    x0 = 0.0
    pef.fill_box(parts, (x0 + 0.5 * dx, 0.5 * dx, 0.0), nx, ny, 1, dx)
    dam_x = x0 + nx * dx
    if with_akinci_walls:
        psi = dx * dx
        n_floor = int(math.ceil(channel_length / dx)) + 4
        xs, ps = [], []
        for i in range(n_floor):
            x = x0 + i * dx
            xs += [(x, -0.5 * dx, 0.0), (x, -1.5 * dx, 0.0)]
        for j in range(ny + 8):
            xs += [(-0.5 * dx, j * dx, 0.0), (-1.5 * dx, j * dx, 0.0)]
        n = len(xs)
        parts._append(np.asarray(xs, float), np.zeros((n, 3)), np.full(n, pef.BOUNDARY, np.uint8),
                      np.full(n, psi), np.zeros((n, 3)))
    return parts, dam_x


def wcsph_dam_break_params(g: float = 9.81) -> "pef.WcsphParams":
    """The parameters wcsph_dam_break_front_at runs with (validation.rs:424-429)."""
    p = pef.WcsphParams.water(1000.0, 50.0, (0.0, -g, 0.0))
    p.adapt_sound_speed = False
    p.xsph_epsilon = 0.0
    p.visc_alpha = 0.0
    return p


def wcsph_dam_break_front_at(dx: float, h0: float, l_res: float, t_end: float, channel_length: float,
                             options: Optional["pef.Options"] = None) -> Tuple[float, float, float]:
    """validation.rs:415-439: (x_front_height, Ritter rel error, t) after
    ceil(t_end / dt) steps of dt = min(0.2 dx / 50, 2e-4)."""
    parts, dam_x = seed_dam_break_particles(dx, h0, l_res, channel_length, 2.2, True, True)
    g = 9.81
    world = pef.WcsphWorld(parts, wcsph_dam_break_params(g), options)
    try:
        dt = min(0.2 * dx / 50.0, 2e-4)
        steps = max(int(math.ceil(t_end / dt)), 1)
        world.step(dt, steps)
        xf = world.front_x_height(dx, h0, channel_length, 0.02)
        t = world.t
        return xf, ritter_front_rel_error(xf, dam_x, g, h0, t), t
    finally:
        world.close()
