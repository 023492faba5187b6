"""Dam-break scenes of the benchmark configurations, built with the library's
own scene functions (fill_box, jitter, fill_box_boundary, GPU Akinci psi).

Every parameter the configurations leave open is fixed here once:
  - walls are one layer of Akinci boundary particles on the tank faces
    (fill_box_boundary + recomputed_akinci_psi), the reference's only WCSPH
    wall model (wcsph.rs:144-149, 244-255);
  - particle mass = mass_calibrated_to_rest(rho0, dx, h) (particles.rs:178-215);
  - the fluid block starts 2 dx from the walls, at (2 dx, 2 dx, 2 dx) in the
    tank corner, as in the reference's 3-D WCSPH bench (benches/fluid.rs:65-96:
    block at 0.1 m, walls at 0, dx 0.05); closer blocks start with wall
    densities of 1.4-1.6 rho0 under a one-layer Akinci wall;
  - the sound speed is fixed (adapt_sound_speed = false), as in every WCSPH
    dam-break and validation run of the reference (validation.rs:430,
    tests/coupling_oracles.rs:73, tests/fluid_analytic.rs:122, 282): with
    c >= 10 v_max the free-surface start-up transient feeds back through
    B = rho0 c^2 / gamma and the f64 step itself diverges within 60 steps;
  - other unstated WCSPH parameters keep WcsphParams::water defaults
    (XSPH 0.25, beta 0, eps 0.01);
  - jitter: per fluid particle in fill order, per component x, y, z,
    uniform(-0.01 dx, 0.01 dx) from xoshiro256** seeded with `seed`.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace
from typing import List, Optional, Tuple

import numpy as np

_M64 = (1 << 64) - 1


class Xoshiro256StarStar:
    """xoshiro256** seeded through SplitMix64, with uniform(lo, hi) from the
    high 53 bits, as the reference's scene generators draw
    (crates/pe-math/src/rng.rs:35-70)."""

    def __init__(self, seed: int):
        sm = seed & _M64
        s = []
        for _ in range(4):
            sm = (sm + 0x9E3779B97F4A7C15) & _M64
            z = sm
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            s.append(z ^ (z >> 31))
        if not any(s):
            s[0] = 0x9E3779B97F4A7C15
        self.s = s

    @staticmethod
    def _rotl(x: int, k: int) -> int:
        return ((x << k) | (x >> (64 - k))) & _M64

    def next_u64(self) -> int:
        s = self.s
        result = (self._rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def next_f64(self) -> float:
        return float(self.next_u64() >> 11) * (1.0 / float(1 << 53))

    def uniform(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.next_f64()


@dataclass(frozen=True)
class DamBreak:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    tank: Tuple[float, float, float]
    dt: float
    steps: int
    jitter: float = 0.01  # fraction of dx
    seed: int = 0
    h_over_dx: float = 2.0
    rest_density: float = 1000.0
    sound_speed: float = 20.0
    visc_alpha: float = 0.05
    gravity: Tuple[float, float, float] = (0.0, 0.0, -9.81)
    dim: int = 3
    wall_gap: float = 2.0  # distance of the first fluid layer from the walls, in dx
    adapt_sound_speed: bool = False
    # static box obstacle (centre, half extents) sampled like FluidRigidCoupler::sample_box
    obstacle: Optional[Tuple[Tuple[float, float, float], Tuple[float, float, float]]] = None

    @property
    def h(self) -> float:
        return self.h_over_dx * self.dx

    @property
    def n_fluid(self) -> int:
        return self.nx * self.ny * self.nz


# The block spans the tank width in y: tank y = (ny - 1 + 2 * wall_gap) dx.
# 8,000-particle CPU case: 20^3 block, dx 0.02, 1.0 x 0.5 x 0.5 m tank, 1,000 steps at 1e-4 s.
CPU_8K = DamBreak("dam_break_8k", 20, 20, 20, 0.02, (1.0, 0.5, 0.5), 1e-4, 1000, jitter=0.0)
# 65,536-particle parity scene: 64 x 32 x 32 at dx 0.005, 200 steps at 2.5e-5 s.
CHECK_64K = DamBreak("dam_break_64k", 64, 32, 32, 0.005, (0.8, 0.175, 0.2), 2.5e-5, 200)
# 1,048,576-particle GPU benchmark: 128 x 64 x 128 at dx 0.005, 500 steps at 2.5e-5 s.
GPU_1M = DamBreak("dam_break_1m", 128, 64, 128, 0.005, (1.6, 0.335, 0.8), 2.5e-5, 500)
# 8,388,608-particle large run: 256 x 128 x 256 at dx 0.005.
GPU_8M = DamBreak("dam_break_8m", 256, 128, 256, 0.005, (3.2, 0.655, 1.6), 2.5e-5, 2000)

# The large hybrid run: the 8M dam break with a static 0.2 x 0.2 x 0.3 m box resting on the floor at
# the tank's x-y centre (12 802 one-sided samples).
HYBRID_8M = DamBreak("dam_break_8m_obstacle", 256, 128, 256, 0.005, (3.2, 0.655, 1.6), 2.5e-5, 2000,
                     obstacle=((1.6, 0.3275, 0.15), (0.1, 0.1, 0.15)))

CONFIGS = {c.name: c for c in (CPU_8K, CHECK_64K, GPU_1M, GPU_8M, HYBRID_8M)}


def _round_half_away(x: float) -> int:
    return int(math.floor(x + 0.5))


def batch_configs(n_scenes: int = 256, seed: int = 1, dx: float = 0.01, dt: float = 5e-5,
                  steps: int = 200) -> List[DamBreak]:
    """Independent dam-break scenes of uneven size for one batched world.

    Per scene, in order, from xoshiro256** seeded with `seed`: a target count
    N = exp(uniform(ln 2048, ln 65536)) (log-uniform) and an aspect
    a = uniform(0.5, 2). The block is nx = round(cbrt(N a)) by ny = nz =
    round(sqrt(N / nx)) particles at spacing dx, h = 2 dx; the realised count
    is nx ny nz. The tank is 2.5 Lx long and 1.25 Lz high (L = n dx, rounded
    up to whole dx), with the block 2 dx from the walls and spanning the tank
    width. No jitter. dt 5e-5 s keeps the 1M dam break's acoustic CFL number
    c dt / h = 0.05 at dx = 0.01 m.
    """
    rng = Xoshiro256StarStar(seed)
    out = []
    gap = 2.0
    for k in range(n_scenes):
        n_target = math.exp(rng.uniform(math.log(2048.0), math.log(65536.0)))
        aspect = rng.uniform(0.5, 2.0)
        nx = max(1, _round_half_away((n_target * aspect) ** (1.0 / 3.0)))
        ny = max(1, _round_half_away(math.sqrt(n_target / nx)))
        nz = ny
        tx = max(math.ceil(2.5 * nx), nx - 1 + 2 * int(gap) + 1)
        tz = max(math.ceil(1.25 * nz), nz - 1 + 2 * int(gap))
        tank = (round(tx * dx, 10), round((ny - 1 + 2 * gap) * dx, 10), round(tz * dx, 10))
        out.append(DamBreak(f"batch_{k}", nx, ny, nz, dx, tank, dt, steps, jitter=0.0))
    return out


BATCH_256 = "batch_256"


def small(nx=12, ny=8, nz=10, dx=0.01, steps=20, jitter=0.01, dim=3) -> DamBreak:
    """A small dam break for quick parity runs (tank 2.5x the block in x)."""
    tank = (round(2.5 * nx * dx, 10), round((ny + 3) * dx, 10), round(1.25 * nz * dx + 3 * dx, 10))
    if dim == 2:
        tank = (tank[0], tank[2], 0.0)
    return DamBreak(f"small_{nx}x{ny}x{nz}", nx, ny, nz, dx, tank, 2.5e-5 * dx / 0.005, steps, jitter=jitter,
                    dim=dim, gravity=(0.0, -9.81, 0.0) if dim == 2 else (0.0, 0.0, -9.81))


def build(cfg: DamBreak):
    """Particles and WcsphParams of a dam break, built through libpef."""
    import pef

    m = pef.mass_calibrated_to_rest(cfg.rest_density, cfg.dx, cfg.h, cfg.dim)
    parts = pef.Particles(cfg.dim, cfg.rest_density, m, cfg.h)
    o = cfg.wall_gap * cfg.dx
    origin = (o, o, o) if cfg.dim == 3 else (o, o, 0.0)
    pef.fill_box(parts, origin, cfg.nx, cfg.ny, cfg.nz, cfg.dx)
    if cfg.jitter:
        pef.jitter_positions(parts, cfg.seed, cfg.jitter * cfg.dx)
    pef.fill_box_boundary(parts, (0.0, 0.0, 0.0), cfg.tank, cfg.dx, 1)
    pef.recomputed_akinci_psi(parts)
    if cfg.obstacle is not None:  # rigid samples keep psi = dx^dim, as the coupler samples them
        sample_box(parts, cfg.obstacle[0], cfg.obstacle[1], cfg.dx)
    params = pef.WcsphParams.water(cfg.rest_density, cfg.sound_speed, cfg.gravity)
    params.visc_alpha = cfg.visc_alpha
    params.adapt_sound_speed = cfg.adapt_sound_speed
    return parts, params


def sample_box(parts, centre, half, dx: float, layers: int = 1):
    """FluidRigidCoupler::sample_box (coupling.rs:103-231) for a body at rest at
    `centre`: one-sided Akinci samples of an axis-aligned box surface with
    outward face normals and psi = dx^dim (RIGID_PSI_SCALE = 1), appended in
    the reference's loop order (z faces, then y faces, then x faces per layer;
    2-D: y faces, then x faces) and placed at centre + local."""
    import pef
    from pef.coupling import box_samples

    xs, ns = box_samples(parts.dim, half, dx, layers)
    psi = dx * dx if parts.dim == pef.D2 else dx * dx * dx
    n = len(xs)
    x = np.asarray(centre, float) + np.asarray(xs, float).reshape(-1, 3)
    parts._append(x, np.zeros((n, 3)), np.full(n, pef.BOUNDARY, np.uint8), np.full(n, psi),
                  np.asarray(ns, float).reshape(-1, 3))
    return n
