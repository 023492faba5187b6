"""Shared scene construction for the parity tests: the same dam-break
configuration built once through libpef (the product) and once through the
oracle, each with its own code."""
import numpy as np

import pef
import pyoracle as O
from pef import scenes


def oracle_scene(cfg):
    """(params dict, arrays dict) of the oracle for a DamBreak config."""
    m = O.mass_calibrated_to_rest(cfg.rest_density, cfg.dx, cfg.h, cfg.dim)
    P = O.Particles(cfg.dim, cfg.rest_density, m, cfg.h)
    o = cfg.wall_gap * cfg.dx
    P.fill_box((o, o, o) if cfg.dim == 3 else (o, o, 0.0), cfg.nx, cfg.ny, cfg.nz, cfg.dx)
    if cfg.jitter:
        P.jitter(cfg.seed, cfg.jitter * cfg.dx)
    P.fill_box_boundary((0.0, 0.0, 0.0), cfg.tank, cfg.dx, 1)
    P.recompute_psi()
    a = P.arrays()
    prm = dict(dim=cfg.dim, rest_density=cfg.rest_density, mass=m, h=cfg.h, sound_speed=cfg.sound_speed,
               visc_alpha=cfg.visc_alpha, gravity=cfg.gravity, adapt_sound_speed=cfg.adapt_sound_speed)
    return prm, a


def oracle_world(cfg, threads=1, **over):
    prm, a = oracle_scene(cfg)
    prm.update(over)
    return O.World(prm, a["x"], a["kind"], psi=a["psi"], threads=threads), a


def gpu_world(cfg, options=None, **over):
    parts, params = scenes.build(cfg)
    for k, v in over.items():
        setattr(params, k, v)
    return pef.WcsphWorld(parts, params, options), parts


def rel_rms(a, b, mask=None):
    a, b = np.asarray(a), np.asarray(b)
    if mask is not None:
        a, b = a[mask], b[mask]
    return float(np.sqrt(np.mean(((a - b) / b) ** 2)))


def vec_rel_rms(a, b, mask=None):
    """RMS of |a - b| over RMS of |b| (vector fields)."""
    a, b = np.asarray(a), np.asarray(b)
    if mask is not None:
        a, b = a[mask], b[mask]
    num = np.sqrt(np.mean(np.sum((a - b) ** 2, axis=1)))
    den = np.sqrt(np.mean(np.sum(b ** 2, axis=1)))
    return float(num / den)
