"""Pins the oracle's grid, neighbour enumeration, scene builders and WCSPH step:
brute-force restatements of the grid by a different method (numpy lexsort and
all-pairs scans), the reference's own WCSPH tests (tests/fluid_analytic.rs),
closed-form particle counts, and bit identity between the reference-structured
STRICT path and the faster FAST / gather paths the large parity runs use.
"""
import math

import numpy as np
import pytest

import pyoracle as O


def brute_grid(x, cell, dim):
    inv = 1.0 / cell
    keys = np.array([[O.floor_i32(p[0] * inv), O.floor_i32(p[1] * inv),
                      0 if dim == 2 else O.floor_i32(p[2] * inv)] for p in x], dtype=np.int64)
    idx = np.arange(len(x))
    order = np.lexsort((idx, keys[:, 2], keys[:, 1], keys[:, 0]))
    sk = keys[order]
    change = np.ones(len(x), bool)
    change[1:] = np.any(sk[1:] != sk[:-1], axis=1)
    starts = np.nonzero(change)[0]
    ends = np.append(starts[1:], len(x))
    return keys, order, sk[starts], np.stack([starts, ends], 1)


def world_from_positions(x, dim=3, h=0.05, kind=None):
    n = len(x)
    kind = np.zeros(n, np.uint8) if kind is None else kind
    prm = dict(dim=dim, rest_density=1000.0, mass=0.01, h=h)
    return O.World(prm, x, kind)


@pytest.mark.parametrize("dim", [2, 3])
def test_grid_matches_bruteforce_sort(dim):
    rng = np.random.default_rng(3)
    x = rng.uniform(-0.37, 0.41, size=(700, 3))
    x[:50] = x[50:100]  # duplicates share cells and keys
    x[100] = [0.1, 0.2, 0.3]  # exact cell faces at cell = 0.1
    x[101] = [-0.1, -0.2, 0.0]
    if dim == 2:
        x[:, 2] = 0.0
    h = 0.05
    W = world_from_positions(x, dim, h)
    W.rebuild_grid()
    g = W.grid()
    keys, order, occ, rng_ = brute_grid(x, 2 * h, dim)
    assert np.array_equal(g["keys"], keys)
    assert np.array_equal(g["order"].astype(np.int64), order)
    assert np.array_equal(g["occupied"], occ)
    assert np.array_equal(g["ranges"].astype(np.int64), rng_)


def test_reference_grid_determinism_case():
    # grid.rs:159-181
    x = np.array([[0.1, 0.1, 0.1], [1.1, 0.1, 0.1], [0.2, 0.1, 0.1], [-0.5, 0.0, 0.0]])
    W = world_from_positions(x, 3, h=0.5)  # cell = 1.0
    W.rebuild_grid()
    nb = list(W.neighbours(0))
    assert 0 in nb and 2 in nb and 1 in nb
    assert nb.index(0) < nb.index(2)
    g1 = W.grid()
    W2 = world_from_positions(x, 3, h=0.5)
    W2.rebuild_grid()
    g2 = W2.grid()
    assert np.array_equal(g1["order"], g2["order"]) and np.array_equal(g1["occupied"], g2["occupied"])


@pytest.mark.parametrize("dim", [2, 3])
def test_neighbours_match_bruteforce(dim):
    rng = np.random.default_rng(11)
    x = rng.uniform(0.0, 0.5, size=(400, 3))
    if dim == 2:
        x[:, 2] = 0.0
    h = 0.04
    W = world_from_positions(x, dim, h)
    W.rebuild_grid()
    keys = W.grid()["keys"].astype(np.int64)
    for i in [0, 7, 133, 399]:
        d = keys - keys[i]
        ok = np.all(np.abs(d) <= 1, axis=1)
        cand = np.nonzero(ok)[0]
        # reference order: dz outer, dy, dx inner; ascending index in a cell
        sk = sorted(cand, key=lambda j: (d[j, 2], d[j, 1], d[j, 0], j))
        assert list(W.neighbours(i).astype(np.int64)) == sk
        # every particle within the support is a candidate
        r = np.linalg.norm(x - x[i], axis=1)
        assert set(np.nonzero(r < 2 * h)[0]) <= set(cand)


def test_fill_counts_closed_form():
    def shell(n):
        nx, ny, nz = n
        return 2 * nx * ny + 2 * nx * (nz - 2) + 2 * (ny - 2) * (nz - 2)
    for (mx, dx, n) in [((1.0, 0.5, 0.5), 0.02, (51, 26, 26)), ((0.6, 0.6, 0.4), 0.05, (13, 13, 9)),
                        ((0.8, 0.16, 0.2), 0.005, (161, 33, 41))]:
        P = O.Particles(3, 1000.0, 1.0, 2 * dx)
        P.fill_box_boundary((0, 0, 0), mx, dx, 1)
        assert len(P) == shell(n)
        a = P.arrays()
        x = a["x"]
        on_face = (np.isclose(x[:, 0], 0) | np.isclose(x[:, 0], mx[0]) | np.isclose(x[:, 1], 0) |
                   np.isclose(x[:, 1], mx[1]) | np.isclose(x[:, 2], 0) | np.isclose(x[:, 2], mx[2]))
        assert on_face.all()
        assert np.all(a["kind"] == O.BOUNDARY)
    assert shell((51, 26, 26)) == 6252 and shell((13, 13, 9)) == 674 and shell((161, 33, 41)) == 25602
    P = O.Particles(3, 1000.0, 1.0, 0.04)
    P.fill_box((0.01, 0.02, 0.03), 3, 4, 5, 0.02)
    x = P.arrays()["x"]
    assert len(x) == 60
    # iz -> iy -> ix order (particles.rs:233-235)
    assert np.allclose(x[1], [0.03, 0.02, 0.03]) and np.allclose(x[3], [0.01, 0.04, 0.03])
    assert np.allclose(x[12], [0.01, 0.02, 0.05])


def test_mass_calibration_gives_rest_density_in_bulk():
    # The calibrated mass makes an interior lattice particle sum to rho0: an
    # independent path (grid + density summation) against particles.rs:178-215.
    dx, h, rho0 = 0.02, 0.04, 1000.0
    m = O.mass_calibrated_to_rest(rho0, dx, h, 3)
    P = O.Particles(3, rho0, m, h)
    P.fill_box((0, 0, 0), 13, 13, 13, dx)
    a = P.arrays()
    W = O.World(dict(dim=3, rest_density=rho0, mass=m, h=h), a["x"], a["kind"])
    W.rebuild_grid()
    W.compute_density()
    rho = W.state()["rho"]
    centre = 6 + 13 * 6 + 169 * 6
    assert abs(rho[centre] - rho0) < 1e-9


def test_psi_grid_form_bit_identical_to_strict():
    P1 = O.Particles(3, 1000.0, 1.0, 0.04)
    P1.fill_box_boundary((0, 0, 0), (0.4, 0.3, 0.2), 0.02, 2)
    P2 = O.Particles(3, 1000.0, 1.0, 0.04)
    P2.fill_box_boundary((0, 0, 0), (0.4, 0.3, 0.2), 0.02, 2)
    P1.recompute_psi(strict=True)
    P2.recompute_psi(strict=False)
    a, b = P1.arrays()["psi"], P2.arrays()["psi"]
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert np.all(a > 0)


def dam_scene(nx=6, ny=5, nz=7, dx=0.02, tank=(0.3, 0.14, 0.2), jitter=0.0):
    h = 2 * dx
    m = O.mass_calibrated_to_rest(1000.0, dx, h, 3)
    P = O.Particles(3, 1000.0, m, h)
    P.fill_box((0.5 * dx,) * 3, nx, ny, nz, dx)
    if jitter:
        P.jitter(0, jitter * dx)
    P.fill_box_boundary((0, 0, 0), tank, dx, 1)
    P.recompute_psi()
    a = P.arrays()
    prm = dict(dim=3, rest_density=1000.0, mass=m, h=h, sound_speed=20.0, visc_alpha=0.05,
               gravity=(0.0, 0.0, -9.81))
    return prm, a


def test_fast_and_gather_paths_bit_identical_to_strict():
    prm, a = dam_scene(jitter=0.01)
    Ws = O.World(prm, a["x"], a["kind"], psi=a["psi"], mode=O.STRICT)
    Wf = O.World(prm, a["x"], a["kind"], psi=a["psi"], mode=O.FAST)
    Wg = O.World(prm, a["x"], a["kind"], psi=a["psi"], mode=O.FAST, threads=4)
    for W in (Ws, Wf, Wg):
        W.step(1e-4, 6)
    ss, sf, sg = Ws.state(), Wf.state(), Wg.state()
    for k in ("x", "v", "a", "rho", "p", "boundary_force"):
        assert np.array_equal(ss[k].view(np.uint64), sf[k].view(np.uint64)), k
        assert np.array_equal(ss[k].view(np.uint64), sg[k].view(np.uint64)), k
    assert Ws.sound_speed == Wf.sound_speed == Wg.sound_speed
    assert Ws.diagnostics()["hash"] == Wg.diagnostics()["hash"]


def test_reference_kat_mass_conserved_wcsph_2d():
    # tests/fluid_analytic.rs:103-137
    dx, rho0 = 0.05, 1000.0
    h = 2.4 * dx
    m = O.mass_from_spacing(rho0, dx, 2)
    P = O.Particles(2, rho0, m, h)
    P.fill_box((0.1, 0.1, 0.0), 6, 6, 1, dx)
    P.fill_box_boundary((0, 0, 0), (0.5, 0.5, 0), dx, 1)
    a = P.arrays()
    prm = dict(dim=2, rest_density=rho0, mass=m, h=h, sound_speed=40.0, gravity=(0.0, -9.80665, 0.0),
               adapt_sound_speed=False, visc_alpha=0.05)
    W = O.World(prm, a["x"], a["kind"], psi=a["psi"])
    m0 = W.diagnostics()["total_mass"]
    W.step(0.0002, 30)
    assert abs(W.diagnostics()["total_mass"] - m0) < 1e-12
    assert np.all(W.state()["x"][:, 2] == 0.0)


def test_reference_kat_pair_forces_conserve_momentum():
    # tests/fluid_analytic.rs:139-169
    dx, rho0 = 0.1, 1000.0
    h = 2.2 * dx
    m = O.mass_from_spacing(rho0, dx, 3)
    x = np.array([[0.0, 0, 0], [dx, 0, 0]])
    v = np.array([[1.0, 0, 0], [-0.5, 0.2, 0]])
    prm = dict(dim=3, rest_density=rho0, mass=m, h=h, sound_speed=50.0, gravity=(0, 0, 0),
               adapt_sound_speed=False, xsph_eps=0.0)
    W = O.World(prm, x, np.zeros(2, np.uint8), v=v)
    W.rebuild_grid()
    W.compute_density()
    W.compute_pressure()
    W.compute_forces()
    fsum = (W.state()["a"] * m).sum(0)
    assert np.linalg.norm(fsum) < 1e-8


def test_momentum_conserved_random_cluster():
    # Antisymmetric pair sums: sum m a = 0 to roundoff for a fluid-only cloud.
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 0.12, size=(300, 3))
    v = rng.normal(0, 0.3, size=(300, 3))
    prm = dict(dim=3, rest_density=1000.0, mass=0.001, h=0.02, sound_speed=20.0, gravity=(0, 0, 0),
               visc_alpha=0.1, visc_beta=0.2)
    W = O.World(prm, x, np.zeros(300, np.uint8), v=v)
    W.rebuild_grid()
    W.compute_density()
    W.compute_pressure()
    W.compute_forces()
    a = W.state()["a"]
    scale = np.abs(a).sum()
    assert np.linalg.norm(a.sum(0)) < 1e-12 * scale


def test_reference_kat_hydrostatic_ordering_2d():
    # tests/fluid_analytic.rs:275-327
    dx, rho0 = 0.05, 1000.0
    hsm = 2.2 * dx
    m = O.mass_from_spacing(rho0, dx, 2)
    P = O.Particles(2, rho0, m, hsm)
    P.fill_box((0.2, 0.15, 0.0), 6, 10, 1, dx)
    P.fill_box_boundary((0.1, 0.05, 0), (0.6, 0.8, 0), dx, 2)
    a = P.arrays()
    prm = dict(dim=2, rest_density=rho0, mass=m, h=hsm, sound_speed=30.0, gravity=(0.0, -9.80665, 0.0),
               adapt_sound_speed=False, xsph_eps=0.2, visc_alpha=0.1)
    W = O.World(prm, a["x"], a["kind"], psi=a["psi"])
    W.step(0.0003, 300)
    W.rebuild_grid()
    W.compute_density()
    W.compute_pressure()
    s = W.state()
    fl = a["kind"] == O.FLUID
    ys = sorted(zip(s["x"][fl, 1], s["p"][fl]))
    n = len(ys)
    lo = np.mean([p for _, p in ys[: n // 3]])
    hi = np.mean([p for _, p in ys[2 * n // 3:]])
    assert lo > hi


def test_reference_kat_mean_density_error_helper():
    # tests/fluid_analytic.rs:363-369
    W = O.World(dict(dim=3, rest_density=1000.0, mass=1.0, h=0.1), np.zeros((1, 3)), np.zeros(1, np.uint8))
    W.set_rho_p(rho=np.array([1010.0]))
    assert abs(W.diagnostics()["mean_density_error"] - 0.01) < 1e-12


def test_adaptive_sound_speed_monotone():
    prm, a = dam_scene()
    W = O.World(prm, a["x"], a["kind"], psi=a["psi"])
    v = np.zeros_like(a["x"])
    v[a["kind"] == O.FLUID] = [3.0, 0.0, 0.0]
    W.set_state(v=v)
    W.step(1e-5, 1)
    assert W.sound_speed >= 30.0 - 1e-9
    W.set_state(v=np.zeros_like(v))
    W.step(1e-5, 1)
    assert W.sound_speed >= 30.0 - 1e-9  # never decreases (wcsph.rs:159-169)
