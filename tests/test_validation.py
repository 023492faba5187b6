"""Host side of the dam-break quantities of interest (pef.validation), pinned
by closed forms of Ritter's solution and by counts of the seeded scene."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import pef
import pyoracle as O
from pef import validation as V

G, H0 = 9.81, 0.4


@pytest.mark.parametrize("t", [0.05, 0.3, 1.7])
def test_ritter_height_closed_forms(t):
    c0 = math.sqrt(G * H0)
    assert V.ritter_c0(G, H0) == c0
    assert V.ritter_front(G, H0, t) == 2.0 * c0 * t
    # h(0, t) = 4/9 h0 for every t > 0 (the dam site keeps critical depth)
    assert V.ritter_height(G, H0, 0.0, t) == pytest.approx(4.0 / 9.0 * H0, rel=1e-14)
    # continuous at the rarefaction head x = -c0 t and at the front x = 2 c0 t
    assert V.ritter_height(G, H0, -c0 * t * (1 - 1e-12), t) == pytest.approx(H0, rel=1e-9)
    assert V.ritter_height(G, H0, 2 * c0 * t * (1 - 1e-9), t) == pytest.approx(0.0, abs=1e-15)
    assert V.ritter_height(G, H0, -c0 * t - 1e-3, t) == H0
    assert V.ritter_height(G, H0, 2 * c0 * t + 1e-3, t) == 0.0
    # volume: what leaves the reservoir equals what lies past the dam
    lost = quad(lambda x: H0 - V.ritter_height(G, H0, x, t), -c0 * t, 0.0, epsabs=1e-14)[0]
    gained = quad(lambda x: V.ritter_height(G, H0, x, t), 0.0, 2 * c0 * t, epsabs=1e-14)[0]
    assert lost == pytest.approx(gained, rel=1e-10)
    # both equal (8/27) h0 c0 t: h = (2 c0 - x/t)^2 / (9 g) integrated over [0, 2 c0 t]
    assert gained == pytest.approx(8.0 / 27.0 * H0 * c0 * t, rel=1e-10)


def test_ritter_height_before_release():
    assert V.ritter_height(G, H0, -0.1, 0.0) == H0
    assert V.ritter_height(G, H0, 0.0, 0.0) == 0.0
    assert V.ritter_height(G, H0, 0.1, -1.0) == 0.0


def test_front_rel_error_and_height_field_front():
    t, dam_x = 0.25, 1.8
    x_ref = dam_x + 2 * math.sqrt(G * H0) * t
    assert V.ritter_front_rel_error(x_ref, dam_x, G, H0, t) == 0.0
    assert V.ritter_front_rel_error(1.1 * x_ref, dam_x, G, H0, t) == pytest.approx(0.1, rel=1e-12)
    hf = [(0.5, 0.3), (1.5, 0.01), (2.5, 0.008), (3.5, 0.0), (4.5, 0.009)]
    assert V.front_x_from_height_field(hf, H0, 0.02) == 4.5  # rightmost wet bin, gaps allowed
    assert V.front_x_from_height_field(hf, H0, 0.05) == 0.5
    assert V.front_x_from_height_field([(0.5, 0.0)], H0) == -math.inf


def test_height_profile_error_skips_dry_bins():
    t, dam_x = 0.2, 1.0
    exact = [(x, V.ritter_height(G, H0, x - dam_x, t)) for x in np.linspace(0.0, 4.0, 81)]
    e, n = V.ritter_height_profile_rel_error(exact, dam_x, G, H0, t)
    assert e == 0.0
    wet = sum(1 for _, h in exact if h >= 0.02 * H0)
    assert n == wet
    off = [(x, h + 0.04) for x, h in exact]  # every bin now wet
    e, n = V.ritter_height_profile_rel_error(off, dam_x, G, H0, t)
    assert n == len(exact) and e == pytest.approx(0.1, rel=1e-12)


def test_rust_round_is_half_away_from_zero():
    assert [V._rust_round(v) for v in (2.5, 3.5, 0.49999999999999994, -1.2, 0.0, 7.0)] == [3, 4, 0, 0, 0, 7]


def test_seed_dam_break_particles_layout():
    dx, h0, l_res, ch = 0.01, 0.1, 0.08, 0.3
    parts, dam_x = V.seed_dam_break_particles(dx, h0, l_res, ch)
    nx, ny = 8, 10
    assert dam_x == nx * dx
    assert parts.fluid_count() == nx * ny
    n_floor = int(math.ceil(ch / dx)) + 4
    assert len(parts) == nx * ny + 2 * n_floor + 2 * (ny + 8)
    assert parts.dim == pef.D2 and parts.h == 2.2 * dx
    assert parts.mass == O.mass_calibrated_to_rest(1000.0, dx, 2.2 * dx, 2)
    fl = parts.kind == pef.FLUID
    assert np.array_equal(np.unique(parts.x[fl, 0]), np.array([0.5 * dx + i * dx for i in range(nx)]))
    b = parts.x[~fl]
    assert np.array_equal(b[:4, :2], [[0.0, -0.5 * dx], [0.0, -1.5 * dx], [dx, -0.5 * dx], [dx, -1.5 * dx]])
    assert np.all(parts.psi[~fl] == dx * dx)
    # small reservoirs are clamped to six particles a side
    p2, d2 = V.seed_dam_break_particles(0.05, 0.1, 0.1, 0.5)
    assert p2.fluid_count() == 36 and d2 == 6 * 0.05


def test_sample_box_matches_the_reference_layout():
    # coupling.rs:103-231: per layer 2 (nx+1)(ny+1) + 2 (nx+1)(nz-1) + 2 (ny-1)(nz-1) samples;
    # the 8M configuration's 0.2 x 0.2 x 0.3 m obstacle box at dx 5 mm has 12 802
    from pef import scenes
    P = pef.Particles(pef.D3, 1000.0, 0.125, 0.01)
    n = scenes.sample_box(P, (1.6, 0.3275, 0.15), (0.1, 0.1, 0.15), 0.005)
    assert n == len(P) == 12802
    x, nrm = P.x, P.n_out
    assert np.allclose(x.min(0), [1.5, 0.2275, 0.0]) and np.allclose(x.max(0), [1.7, 0.4275, 0.3])
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0)
    # every sample lies on the face its normal points out of
    c, h = np.array([1.6, 0.3275, 0.15]), np.array([0.1, 0.1, 0.15])
    a = np.abs(nrm).argmax(1)
    face = c[a] + np.sign(nrm[np.arange(n), a]) * h[a]
    assert np.allclose(x[np.arange(n), a], face)
    assert np.all(P.psi == 0.005 ** 3) and np.all(P.kind == pef.BOUNDARY)
    # first samples: the z faces of i = 0, j = 0, bottom then top
    assert np.allclose(x[:2], [[1.5, 0.2275, 0.0], [1.5, 0.2275, 0.3]])
    Q = pef.Particles(pef.D2, 1000.0, 0.025, 0.01)
    assert scenes.sample_box(Q, (0.5, 0.5, 0.0), (0.1, 0.05, 0.0), 0.01) == 2 * 21 + 2 * 9
