"""The coupled fluid-rigid pieces of the oracle and the host coupling module.

Pins:
  - the oracle's restatement of the rigid world (symplectic Euler with the
    implicit-midpoint rotation drift, gravity and external loads), the
    FluidRigidCoupler (2-D box sampling, sync, reduce), carve_fluid_box and the
    DFSPH pressure correction with its boundary reactions, velocity cap and
    free-slip planes reproduce the reference's locked coupling hash
    0xe9a6a84fef4ce9d5 (docs/DETERMINISM.md:668-682, coupling.rs:739-767) bit
    for bit;
  - closed forms: inertia KATs of the reference's own tests (body.rs:268-284),
    free fall under symplectic Euler, a sphere's torque-free spin turning by
    |w| t about its axis with L unchanged;
  - the host coupling module (pef.coupling: the caller's rigid world, box and
    sphere sampling, carving) is the oracle bit for bit.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from pef import coupling as CP

COUPLING_HASH = 0xE9A6A84FEF4CE9D5


def test_coupling_hash_matches_the_reference_lock():
    h, rigid = O.coupling_tank_hash(80)
    assert h == COUPLING_HASH, hex(h)
    # the hash moves with the run: one step fewer is another value
    assert O.coupling_tank_hash(79)[0] != COUPLING_HASH
    assert np.all(np.isfinite(rigid)) and abs(np.linalg.norm(rigid[3:7]) - 1.0) < 1e-15


def test_dfsph_tank_hash_unchanged_by_the_coupling_terms():
    # velocity cap, reactions and free-slip planes are off in the tank scenario
    assert O.dfsph_tank_hash(40) == 0xF1F2825B5EEAAB0D


def test_inertia_kats():
    # body.rs:276-283: a cube of side 1 and mass 6 has I = m a^2 / 6 = 1
    for I in (O.inertia_solid_box(6.0, 1.0, 1.0, 1.0), np.array(CP.inertia_solid_box(6.0, 1.0, 1.0, 1.0))):
        assert np.allclose(np.diag(I), 1.0, atol=1e-15) and np.count_nonzero(I - np.diag(np.diag(I))) == 0
    I = O.inertia_solid_box(2.0, 0.24, 0.16, 1.0)
    assert np.array_equal(I, np.array(CP.inertia_solid_box(2.0, 0.24, 0.16, 1.0)))
    assert I[2, 2] == pytest.approx(2.0 * (0.24 ** 2 + 0.16 ** 2) / 12.0, rel=1e-15)
    S = O.inertia_solid_sphere(3.0, 0.5)
    assert np.array_equal(S, np.array(CP.inertia_solid_sphere(3.0, 0.5)))
    assert S[0, 0] == pytest.approx(0.4 * 3.0 * 0.25, rel=1e-15)


def _body(mass, inertia, x, q=(1.0, 0.0, 0.0, 0.0), v=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0)):
    return dict(mass=mass, inertia=np.asarray(inertia, float).reshape(9), x=x, q=q, v=v, omega=omega)


def test_free_fall_closed_form():
    g = np.array([0.0, -9.80665, 0.0])
    x0, v0, dt, k = np.array([0.1, 2.0, 0.0]), np.array([0.5, 1.0, 0.0]), 1e-3, 500
    R = O.RigidWorld([_body(2.0, np.eye(3), x0, v=v0)], g)
    for _ in range(k):
        R.step(dt)
    s = R.pack()
    # symplectic Euler: v_k = v0 + k g dt, x_k = x0 + k dt v0 + dt^2 g k (k + 1) / 2
    assert np.allclose(s[7:10], v0 + k * dt * g, rtol=1e-12, atol=1e-12)
    assert np.allclose(s[0:3], x0 + k * dt * v0 + dt * dt * g * k * (k + 1) / 2, rtol=1e-11, atol=1e-12)
    assert np.array_equal(s[3:7], [1.0, 0.0, 0.0, 0.0])


def test_sphere_spin_turns_by_omega_t():
    w = np.array([0.3, -1.2, 2.0])
    R = O.RigidWorld([_body(1.0, O.inertia_solid_sphere(1.0, 0.2), np.zeros(3), omega=w)], np.zeros(3))
    l0 = R.pack()[10:13].copy()
    dt, k = 1e-3, 700
    for _ in range(k):
        R.step(dt)
    s = R.pack()
    th = np.linalg.norm(w) * k * dt
    want = np.concatenate([[math.cos(th / 2)], math.sin(th / 2) * w / np.linalg.norm(w)])
    assert np.allclose(s[3:7], want, atol=1e-12)
    assert np.array_equal(s[10:13], l0)  # no torque: L is untouched
    assert np.allclose(R.omega()[0], w, rtol=1e-12)


def test_host_rigid_world_is_the_oracle_bit_for_bit():
    rng = np.random.default_rng(7)
    bodies = []
    for b in range(3):
        A = rng.normal(size=(3, 3))
        inertia = A @ A.T + 3.0 * np.eye(3)
        inertia = 0.5 * (inertia + inertia.T)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        bodies.append(_body(1.0 + b, inertia, rng.normal(size=3), tuple(q), tuple(rng.normal(size=3)),
                            tuple(rng.normal(size=3))))
    g = (0.0, -9.80665, 0.0)
    R = O.RigidWorld(bodies, g)
    H = CP.RigidWorld([CP.RigidBody(b["mass"]).with_position(b["x"]).with_velocity(b["v"])
                       .with_orientation(b["q"]).with_angular_velocity(b["omega"])
                       .with_inertia(b["inertia"].reshape(3, 3)) for b in bodies], g)
    # with_orientation normalises, as RigidBody::with_orientation does, and World::new once more
    assert np.array_equal(H.pack(), R.pack())
    for k in range(200):
        f, t = rng.normal(size=(3, 3)) * 10.0, rng.normal(size=(3, 3))
        R.step(1e-3, f, t)
        H.set_loads(f, t)
        H.step(1e-3)
        H.clear_loads()
    assert np.array_equal(H.pack(), R.pack())
    assert np.array_equal(np.array(H.omega()), R.omega())


@pytest.mark.parametrize("dim,half,layers", [(2, (0.12, 0.08, 0.0), 1), (2, (0.1, 0.05, 0.0), 2),
                                             (3, (0.1, 0.1, 0.15), 1), (3, (0.05, 0.07, 0.03), 2)])
def test_box_samples_match_the_oracle(dim, half, layers):
    dx = 0.01 if dim == 3 else 0.05 / 2
    P = O.Particles(dim, 1000.0, 1.0, 2.2 * dx)
    C = O.Coupler(1)
    C.sample_box(P, 0, half, dx, layers)
    particle, body, local, local_n = C.samples()
    xs, ns = CP.box_samples(dim, half, dx, layers)
    assert np.array_equal(np.array(xs), local) and np.array_equal(np.array(ns), local_n)
    assert np.array_equal(particle, np.arange(len(xs))) and not body.any()
    a = P.arrays()
    assert np.array_equal(a["x"], local) and np.all(a["kind"] == 1)
    assert np.all(a["psi"] == (dx * dx if dim == 2 else dx ** 3))
    nx, ny = round(2 * half[0] / dx), round(2 * half[1] / dx)
    if dim == 2:
        assert len(xs) == layers * (2 * (nx + 1) + 2 * (ny - 1))
    else:
        nz = round(2 * half[2] / dx)
        assert len(xs) == layers * (2 * (nx + 1) * (ny + 1) + 2 * (nx + 1) * (nz - 1) + 2 * (ny - 1) * (nz - 1))


@pytest.mark.parametrize("dim,radius,dx", [(2, 0.12, 0.02), (2, 0.01, 0.05), (3, 0.12, 0.055), (3, 0.02, 0.05)])
def test_sphere_samples_match_the_oracle(dim, radius, dx):
    P = O.Particles(dim, 1000.0, 1.0, 2.2 * dx)
    C = O.Coupler(1)
    C.sample_sphere(P, 0, radius, dx)
    _, _, local, local_n = C.samples()
    xs, ns = CP.sphere_samples(dim, radius, dx)
    assert np.array_equal(np.array(xs), local) and np.array_equal(np.array(ns), local_n)
    assert np.allclose(np.linalg.norm(local, axis=1), radius, rtol=1e-14)
    assert np.allclose(np.linalg.norm(local_n, axis=1), 1.0, rtol=1e-14)
    n_min = 8 if dim == 2 else 12
    want = (round(2 * math.pi * radius / dx) if dim == 2 else round(4 * math.pi * radius ** 2 / dx ** 2))
    assert len(xs) == max(want, n_min)


def test_carve_matches_the_oracle():
    import pef

    dx = 0.05
    P = O.Particles(2, 1000.0, 1.0, 2.2 * dx)
    P.fill_box((0.5 * dx, 0.5 * dx, 0.0), 18, 8, 1, dx)
    P.fill_box_boundary((0.0, 0.0, 0.0), (0.9, 1.0, 0.0), dx, 2)
    a = P.arrays()
    H = pef.Particles(2, 1000.0, 1.0, 2.2 * dx)
    H._append(a["x"], a["v"], a["kind"], a["psi"], a["n_out"])
    c, half = (0.45, 0.33, 0.0), (0.12, 0.08, 0.0)
    O.carve_fluid_box(P, c, half, dx)
    removed = CP.carve_fluid_box(H, c, half, dx)
    b = P.arrays()
    assert removed == len(a["x"]) - len(b["x"]) > 0
    for k in ("x", "kind", "psi", "n_out"):
        assert np.array_equal(getattr(H, k), b[k])


def test_heave_helpers_closed_forms():
    T = 0.8
    t = np.arange(0.0, 4.0, 2.5e-4)
    ys = list(zip(t, 0.3 + 0.01 * np.sin(2 * np.pi * t / T + 0.3)))
    assert CP.estimate_heave_period(ys) == pytest.approx(T, rel=1e-6)
    assert CP.estimate_heave_period(ys[:7]) is None
    # the ensemble's spectral peak: records of one period with random phases,
    # drifts and noise, on the times the heave test samples (0.1 s to 4 s)
    rng = np.random.default_rng(3)
    tt = np.arange(0.1, 4.0, 2.5e-4)
    runs = [list(zip(tt, 0.4 + 0.02 * rng.normal() * tt + 0.01 * np.sin(2 * np.pi * tt / T + rng.uniform(0, 6.3))
                     + 0.004 * rng.normal(size=len(tt)))) for _ in range(8)]
    # bin spacing of the 8x padded spectrum near 1 / T: T^2 / (8 * 3.9 s)
    assert CP.ensemble_heave_period(runs) == pytest.approx(T, abs=T * T / (8 * 3.9))
    assert CP.ensemble_heave_period(runs[:1]) == pytest.approx(T, abs=T * T / (8 * 3.9))
    with pytest.raises(ValueError):
        CP.ensemble_heave_period([runs[0], runs[1][:-5]])
    assert CP.archimedes_box_draft(500.0, 1000.0, 0.16) == pytest.approx(0.08)
    m, a_wp = 19.2, 0.24
    m_a = CP.barge_heave_added_mass_2d(1000.0, a_wp)
    assert m_a == pytest.approx(1000.0 * math.pi * a_wp ** 2 / 8)
    assert CP.heave_period_with_added_mass(m, 0.0, 1000.0, a_wp, 9.80665) == pytest.approx(
        CP.heave_period_analytic(m, 1000.0, a_wp, 9.80665))
    assert CP.sphere_freefall_accel_added_mass(2200.0, 1000.0, 9.80665) == pytest.approx(
        9.80665 * 1200.0 / 2700.0)
