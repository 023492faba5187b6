"""The coupler's side of the WCSPH step on the device (coupling.rs:267-331).

Bars:
  - sync_from_rigid places every sample at x + R(q) local with normal
    R(q) local_n and velocity v + omega x (R(q) local) bit for bit as the
    reference's FP64 operation order gives it (restated here with numpy, which
    rounds every operation like the device without contraction; velocities
    as their FP32 rounding); rotations keep every |x_s - x_b| = |local| to
    1 ulp scale, and the sample velocities are a rigid field:
    (v_s - v_b) . (x_s - x_b) = 0;
  - reduce_with_state equals the FP64 sums of the device's own boundary_force
    over each body's samples, to 1e-12 relative, also after list rebuilds
    permuted the slots, and per body;
  - a translating one-sided obstacle driven through sync_rigid follows the f64
    oracle driven through set_state: density RMS relative <= 1e-5 every five
    steps, boundary reactions and the reduced force within 1e-4.
"""
import numpy as np
import pytest

import pef
import pyoracle as O
from pef import scenes
from helpers import rel_rms, vec_rel_rms
from test_gpu_parity import _box_surface

pytestmark = pytest.mark.gpu

CFG = scenes.small(16, 8, 10, dx=0.01, steps=30)


def _quat(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([[np.cos(angle / 2)], np.sin(angle / 2) * axis])


def _rotate(q, v):
    """Quat::rotate (quat.rs:102-138) in numpy, the reference's operation order."""
    w, x, y, z = q
    R = [[1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y)],
         [2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x)],
         [2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)]]
    v = np.asarray(v, float)
    return np.stack([R[a][0] * v[:, 0] + R[a][1] * v[:, 1] + R[a][2] * v[:, 2] for a in range(3)], 1)


def _scene():
    """The small dam break with a one-sided box obstacle; (params, particles,
    obstacle ids, obstacle centre, local offsets, local normals)."""
    dx, h = CFG.dx, CFG.h
    m = O.mass_calibrated_to_rest(CFG.rest_density, dx, h, 3)
    o = CFG.wall_gap * dx
    P = O.Particles(3, CFG.rest_density, m, h)
    P.fill_box((o, o, o), CFG.nx, CFG.ny, CFG.nz, dx)
    P.jitter(0, 0.01 * dx)
    P.fill_box_boundary((0.0, 0.0, 0.0), CFG.tank, dx, 1)
    a = P.arrays()
    lo, hi = np.array([0.19, 0.03, 0.03]), np.array([0.23, 0.07, 0.07])
    obs_x, obs_n = _box_surface(lo, hi, dx)
    x = np.concatenate([a["x"], obs_x])
    kind = np.concatenate([a["kind"], np.ones(len(obs_x), np.uint8)])
    n_out = np.concatenate([np.zeros_like(a["x"]), obs_n])
    psi = np.concatenate([a["psi"], np.full(len(obs_x), dx ** 3)])
    G = pef.Particles(3, CFG.rest_density, m, h)
    G._append(x, np.zeros_like(x), kind, psi, n_out)
    pef.recomputed_akinci_psi(G)
    params = pef.WcsphParams.water(CFG.rest_density, CFG.sound_speed, CFG.gravity)
    params.visc_alpha = CFG.visc_alpha
    params.adapt_sound_speed = False
    ids = np.arange(len(a["x"]), len(x))
    c = 0.5 * (lo + hi)
    return params, G, ids, c, obs_x - c, obs_n


def _world():
    params, G, ids, c, local, local_n = _scene()
    W = pef.WcsphWorld(G, params)
    W.set_rigid_samples(ids, np.zeros(len(ids), np.uint32), local, local_n, 1)
    return W, G, ids, c, local, local_n


def _cross(a, b):
    """Vec3::cross (vec3.rs:87-93), row-wise, the reference's operation order."""
    return np.stack([a[..., 1] * b[:, 2] - a[..., 2] * b[:, 1], a[..., 2] * b[:, 0] - a[..., 0] * b[:, 2],
                     a[..., 0] * b[:, 1] - a[..., 1] * b[:, 0]], 1)


def test_sync_places_samples_like_the_coupler():
    W, G, ids, c, local, local_n = _world()
    x_b = c + np.array([0.004, -0.002, 0.003])
    q = _quat([0.3, -0.5, 1.0], 0.4)
    v_b, w_b = np.array([0.3, -0.1, 0.02]), np.array([1.5, -4.0, 2.5])
    W.sync_rigid(x_b[None], q[None], v_b[None], w_b[None])
    xs = W.positions()[ids]
    r = _rotate(q, local)
    assert np.array_equal(xs.view(np.uint64), (x_b + r).view(np.uint64))
    assert np.array_equal(W.normals()[ids], _rotate(q, local_n))
    vs = W.velocities()[ids]
    assert np.array_equal(vs, (v_b + _cross(w_b, r)).astype(np.float32).astype(np.float64))
    assert np.abs(np.sum((vs - v_b) * r, 1)).max() <= 1e-6 * np.abs(vs - v_b).max() * np.abs(r).max()
    d = np.linalg.norm(xs - x_b, axis=1)
    assert np.allclose(d, np.linalg.norm(local, axis=1), rtol=1e-15, atol=1e-17)
    # the identity pose at rest puts the samples back where they were built
    W.sync_rigid(c[None], np.array([[1.0, 0.0, 0.0, 0.0]]))
    assert np.array_equal(W.positions()[ids], c + local)
    assert not W.velocities()[ids].any()
    # the rest of the world is untouched
    others = np.setdiff1d(np.arange(W.n), ids)
    assert np.array_equal(W.positions()[others], G.x[others])
    assert not W.velocities()[others].any()
    assert np.array_equal(W.normals(), G.n_out) and np.array_equal(W.psi(), G.psi)


def test_set_state_keeps_the_turned_normals():
    """A host state write after sync_rigid (a caller mixing both) keeps the
    normals the device turned, and the step then sees the turned samples."""
    W, G, ids, c, local, local_n = _world()
    q = _quat([0.0, 0.0, 1.0], 0.7)
    W.sync_rigid(c[None], q[None])
    turned = W.normals()
    v = W.velocities()
    v[G.kind == pef.FLUID] += [0.01, 0.0, 0.0]
    W.set_state(v=v)
    assert np.array_equal(W.normals(), turned)
    assert np.array_equal(W.velocities(), v.astype(np.float32).astype(np.float64))


def test_reduce_sums_the_reactions_per_body():
    W, G, ids, c, local, local_n = _world()
    body = (np.arange(len(ids)) % 2).astype(np.uint32)  # interleaved: both bodies touch the fluid
    W.set_rigid_samples(ids, body, local, local_n, 2)
    vel = np.array([-0.2, 0.0, 0.05])
    for k in range(CFG.steps):
        W.sync_rigid((c + (k + 1) * CFG.dt * vel)[None].repeat(2, 0), np.tile([1.0, 0.0, 0.0, 0.0], (2, 1)))
        W.step(CFG.dt)
    _, rebuilds = W.list_stats()
    assert rebuilds > 1  # slots were permuted between force evaluations
    com = np.array([c + CFG.steps * CFG.dt * vel, c])
    f, t = W.reduce_rigid(com)
    bf, x = W.boundary_force(), W.positions()
    for b, sel in enumerate((ids[0::2], ids[1::2])):
        fw = bf[sel].sum(0)
        tw = np.cross(x[sel] - com[b], bf[sel]).sum(0)
        assert np.abs(fw).sum() > 0
        assert np.allclose(f[b], fw, rtol=1e-12, atol=1e-12 * np.abs(fw).max())
        assert np.allclose(t[b], tw, rtol=1e-12, atol=1e-12 * np.abs(tw).max())


def test_translating_obstacle_through_sync_matches_oracle():
    W, G, ids, c, local, local_n = _world()
    prm = dict(dim=3, rest_density=CFG.rest_density, mass=G.mass, h=G.h, sound_speed=CFG.sound_speed,
               visc_alpha=CFG.visc_alpha, gravity=CFG.gravity, adapt_sound_speed=False)
    Wo = O.World(prm, G.x, G.kind, psi=G.psi, n_out=G.n_out, threads=4)
    vel = np.array([-0.2, 0.0, 0.05])
    fl = G.kind == pef.FLUID
    for k in range(CFG.steps):
        x_b = c + (k + 1) * CFG.dt * vel
        W.sync_rigid(x_b[None], np.array([[1.0, 0.0, 0.0, 0.0]]), vel[None])
        so = Wo.state()
        xo, vo = so["x"].reshape(-1, 3).copy(), so["v"].reshape(-1, 3).copy()
        xo[ids], vo[ids] = x_b + local, vel
        Wo.set_state(x=xo, v=vo)
        W.step(CFG.dt)
        Wo.step(CFG.dt)
        if k % 5 == 4:
            e = rel_rms(W.densities()[fl], Wo.state()["rho"][fl])
            assert e <= 1e-5, (k, e)
    assert np.array_equal(W.velocities()[ids], np.tile(vel.astype(np.float32), (len(ids), 1)).astype(np.float64))
    bf, bfo = W.boundary_force(), Wo.state()["boundary_force"].reshape(-1, 3)
    assert vec_rel_rms(bf[ids], bfo[ids]) <= 1e-4
    f, _ = W.reduce_rigid((c + CFG.steps * CFG.dt * vel)[None])
    fo = bfo[ids].sum(0)
    assert np.linalg.norm(f[0] - fo) <= 1e-4 * np.linalg.norm(fo)


def test_rigid_sample_errors():
    W, G, ids, c, local, local_n = _world()
    with pytest.raises(pef.PefError):  # a fluid particle cannot be a sample
        W.set_rigid_samples([0], [0], local[:1], local_n[:1], 1)
    with pytest.raises(pef.PefError):
        W.sync_rigid(np.zeros((2, 3)), np.tile([1.0, 0, 0, 0], (2, 1)))  # body count differs
    with pytest.raises(pef.PefError):
        W.sync_rigid(np.array([[np.nan, 0, 0]]), np.array([[1.0, 0, 0, 0]]))
    with pytest.raises(pef.PefError) as e:  # one boundary particle sampled twice
        W.set_rigid_samples([ids[0], ids[0]], [0, 0], local[:2], local_n[:2], 1)
    assert e.value.code == pef.ERR_ARG


def test_one_sided_visibility_follows_particles_between_list_builds():
    """A fluid block moving at 0.01 h per step toward the back of a one-sided
    plate (samples at y = 0, normals +y, boundary_sees_fluid: y_f > -0.15 h):
    its top layer crosses the visibility plane after 8 steps, between the list
    builds the skin asks for after steps 5 and 10. Densities follow the oracle,
    which evaluates visibility for every pair at every step."""
    dx, h = 0.01, 0.02
    m = O.mass_calibrated_to_rest(1000.0, dx, h, 3)
    xs = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3) * dx
    fluid = xs + np.array([0.005, -0.0745, 0.005])  # top layer at y = -0.0045
    gx, gz = np.meshgrid(np.arange(-2, 11), np.arange(-2, 11), indexing="ij")
    plate = np.stack([gx.reshape(-1) * dx, np.zeros(gx.size), gz.reshape(-1) * dx], 1)
    x = np.concatenate([fluid, plate])
    kind = np.concatenate([np.zeros(len(fluid), np.uint8), np.ones(len(plate), np.uint8)])
    v = np.zeros_like(x)
    v[: len(fluid), 1] = 2.0
    psi = np.concatenate([np.zeros(len(fluid)), np.full(len(plate), dx ** 3)])
    n_out = np.zeros_like(x)
    n_out[len(fluid):, 1] = 1.0
    G = pef.Particles(3, 1000.0, m, h)
    G._append(x, v, kind, psi, n_out)
    params = pef.WcsphParams.water(1000.0, 40.0, (0.0, 0.0, 0.0))
    params.visc_alpha = 0.0
    params.xsph_epsilon = 0.0
    params.adapt_sound_speed = False
    W = pef.WcsphWorld(G, params)
    prm = dict(dim=3, rest_density=1000.0, mass=m, h=h, sound_speed=40.0, visc_alpha=0.0, xsph_eps=0.0,
               gravity=(0.0, 0.0, 0.0), adapt_sound_speed=False)
    Wo = O.World(prm, x, kind, v=v, psi=psi, n_out=n_out, threads=4)
    fl = kind == 0
    dt = 1e-4
    crossed = False
    for k in range(14):
        W.step(dt)
        Wo.step(dt)
        rho, rho_o = W.densities()[fl], Wo.state()["rho"][fl]
        assert np.abs(rho / rho_o - 1).max() <= 1e-4, (k, np.abs(rho / rho_o - 1).max())
        crossed = crossed or (Wo.state()["x"].reshape(-1, 3)[fl, 1] > -0.15 * h).any()
    assert crossed
