"""Pins the oracle's closed-form pieces (cubic spline, Tait EOS, Monaghan
viscosity, floor, hashes, RNG, lattice mass) against mathematics that does not
come from the oracle's own code: exact rational arithmetic of the textbook
formulas, exact normalisation integrals, finite differences, the reference's
own measured numbers (docs/FLUID.md:34-40, 110-122) and published test vectors.
"""
import math
from decimal import Decimal, getcontext
from fractions import Fraction as F

import numpy as np
import pytest

import pyoracle as O

getcontext().prec = 60
# pi to 60 digits (Machin), independent of the oracle's constant.
PI = Decimal("3.14159265358979323846264338327950288419716939937510582097494")


def monaghan_poly(q: F) -> F:
    """Monaghan & Lattanzio (1985) cubic spline shape, support 2h."""
    if q < 1:
        return 1 - F(3, 2) * q * q + F(3, 4) * q * q * q
    if q < 2:
        return F(1, 4) * (2 - q) ** 3
    return F(0)


def w_exact(r: float, h: float, dim: int) -> Decimal:
    # q is taken as the IEEE-rounded r/h (one correctly rounded division), so the
    # comparison isolates the evaluation of the polynomial and the normalisation.
    q = F(r / h)
    poly = monaghan_poly(q)
    hh = Decimal(F(h).numerator) / Decimal(F(h).denominator)
    pd = Decimal(poly.numerator) / Decimal(poly.denominator)
    if dim == 3:
        return pd / (PI * hh ** 3)
    return Decimal(10) * pd / (Decimal(7) * PI * hh ** 2)


@pytest.mark.parametrize("dim", [2, 3])
def test_cubic_w_matches_exact_textbook_value(dim):
    h = 0.0123
    rs = [0.0, 1e-9, 0.3 * h, 0.999 * h, h, 1.0000001 * h, 1.5 * h, 1.999999 * h, 2 * h, 2.5 * h]
    for r in rs:
        got = O.w(r, h, dim)
        ex = w_exact(r, h, dim)
        tol = Decimal(8) * Decimal(2.0 ** -52) * abs(ex) + Decimal(1e-300)
        assert abs(Decimal(got) - ex) <= tol, (r, got, ex)
    assert O.w(-1e-3, h, dim) == 0.0
    assert O.w(0.1, 0.0, dim) == 0.0


@pytest.mark.parametrize("dim", [2, 3])
def test_cubic_normalisation_is_exactly_one(dim):
    # Exact polynomial integrals of the shape: 3-D 4*pi*alpha*h^3 * int f q^2 dq,
    # 2-D 2*pi*alpha*h^2 * int f q dq. With alpha from the oracle (h = 1):
    a = O.w(0.0, 1.0, dim)  # alpha, since f(0) = 1
    def integ(pw):
        # int_0^1 (1 - 1.5 q^2 + .75 q^3) q^pw dq + int_1^2 .25 (2-q)^3 q^pw dq, exactly
        first = F(1, pw + 1) - F(3, 2) * F(1, pw + 3) + F(3, 4) * F(1, pw + 4)
        # expand (2-q)^3 = 8 - 12 q + 6 q^2 - q^3
        coeffs = [8, -12, 6, -1]
        second = sum(F(c) * (F(2) ** (k + pw + 1) - 1) / (k + pw + 1) for k, c in enumerate(coeffs)) / 4
        return first + second
    if dim == 3:
        total = 4 * math.pi * a * float(integ(2))
    else:
        total = 2 * math.pi * a * float(integ(1))
    assert abs(total - 1.0) < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_dwdr_is_derivative_of_w(dim):
    # Exact central difference of the exact textbook W (error O(delta^2)).
    h = 0.05
    delta = F(1, 10 ** 18)
    for qf in [0.1, 0.5, 0.9, 1.2, 1.7, 1.95]:
        r = F(qf * h)
        num = (monaghan_poly((r + delta) / F(h)) - monaghan_poly((r - delta) / F(h))) / (2 * delta)
        alpha = O.w(0.0, h, dim)
        ref = float(num) * alpha
        got = O.dwdr(float(r), h, dim)
        assert abs(got - ref) <= 1e-12 * abs(ref), (qf, got, ref)
    assert O.dwdr(0.0, h, dim) == 0.0
    assert O.dwdr(2 * h, h, dim) == 0.0


def test_grad_w_is_gradient_wrt_xi():
    # grad_w(rij) returns grad_i W(|x_j - x_i|): check against a finite difference
    # of W in x_i (kernels.rs:52-75 states the convention).
    h = 0.04
    rij = np.array([0.013, -0.021, 0.007])
    g = O.grad_w(rij, h, 3)
    eps = 1e-9
    fd = np.zeros(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = eps
        # moving x_i by +e changes rij by -e
        fd[k] = (O.w(np.linalg.norm(rij - e), h, 3) - O.w(np.linalg.norm(rij + e), h, 3)) / (2 * eps)
    assert np.allclose(g, fd, rtol=1e-5, atol=0)
    assert np.all(O.grad_w([0.0, 0.0, 0.0], h, 3) == 0.0)
    # Points toward the neighbour for a falling kernel.
    assert np.dot(g, rij) > 0


# --- reference KATs restated over the oracle (kernels.rs:212-330) -----------------
def partition_of_unity_error(h, dim, n_half):
    support = 2.0 * h
    dx = support / n_half
    n = 2 * n_half
    xs = -support + np.arange(n + 1) * dx
    s = 0.0
    if dim == 3:
        for x in xs:
            for y in xs:
                for z in xs:
                    s += O.w(math.sqrt(x * x + y * y + z * z), h, 3) * dx * dx * dx
    else:
        for x in xs:
            for y in xs:
                s += O.w(math.sqrt(x * x + y * y), h, 2) * dx * dx
    return abs(s - 1.0)


def lattice_grad_sums(h, dim, dx):
    support = 2.0 * h
    n = int(math.ceil(support / dx)) + 1
    gsum = np.zeros(3)
    m = np.zeros((3, 3))
    vol = dx * dx if dim == 2 else dx * dx * dx
    zr = [0] if dim == 2 else range(-n, n + 1)
    for ix in range(-n, n + 1):
        for iy in range(-n, n + 1):
            for iz in zr:
                if ix == 0 and iy == 0 and iz == 0:
                    continue
                rij = np.array([ix * dx, iy * dx, iz * dx])
                if np.linalg.norm(rij) >= support:
                    continue
                g = O.grad_w(rij, h, dim)
                gsum += g * vol
                m += vol * np.outer(rij, g)
    dn = dim
    err = max(abs(m[a, b] - (1.0 if a == b else 0.0)) for a in range(dn) for b in range(dn))
    return np.linalg.norm(gsum), err


def test_reference_kat_partition_of_unity_cubic_3d():
    # tests/fluid_analytic.rs:25-35: < 1% at n_half = 24.
    assert partition_of_unity_error(0.1, 3, 24) < 0.01


def test_reference_kat_gradient_consistency_numbers():
    # tests/fluid_analytic.rs:46-59 and docs/FLUID.md:34-40 (measured 1.51e-4, 2.5e-7).
    h = 0.12
    g8, e8 = lattice_grad_sums(h, 2, h / 8)
    g4, e4 = lattice_grad_sums(h, 2, h / 4)
    assert g8 < 1e-12
    assert e4 < 5e-4 and e8 < 1e-5 and e8 < e4
    assert abs(e4 - 1.51e-4) < 0.01e-4
    assert abs(e8 - 2.5e-7) < 0.1e-7


# --- Tait EOS (eos.rs:59-81) ------------------------------------------------------
def test_tait_closed_form_exact():
    rho0, c, gamma = 1000.0, 20.0, 7.0
    for rho in [1000.0, 1000.5, 1010.0, 990.0, 1234.5]:
        got = O.tait_pressure(rho, rho0, c, gamma)
        ratio = F(rho) / F(rho0)
        ex = F(rho0) * F(c) * F(c) / F(gamma) * (ratio ** 7 - 1)
        tol = 1e-12 * abs(float(F(rho0) * F(c) * F(c) / 7)) * 1e-3 + 1e-300
        assert abs(got - float(ex)) <= max(tol, 1e-13 * abs(float(ex))), (rho, got, float(ex))
    assert O.tait_pressure(rho0, rho0, c, gamma) == 0.0  # tests/fluid_analytic.rs:61-70
    assert O.tait_pressure(1010.0, rho0, c, gamma) > 0.0


def test_reference_kat_tait_hydrostatic_profile():
    # validation.rs:513-541 restated; docs/FLUID.md:110-122: max 0.81 %, mean 0.41 % at c0 = 60.
    g, rho0, c0, H, ns = 9.80665, 1000.0, 60.0, 1.0, 40
    rels = []
    for k in range(ns):
        z = 0.0 + (k + 0.5) / ns * (H - 0.0)
        depth = H - z
        rho = rho0 * (1.0 + g * depth / (c0 * c0))
        p = O.tait_pressure(rho, rho0, c0, 7.0)
        p_ref = rho0 * g * depth
        rels.append(abs(p - p_ref) / p_ref)
    assert len(rels) >= 30
    assert max(rels) < 0.01
    assert abs(max(rels) - 0.0081) < 0.0001
    assert abs(np.mean(rels) - 0.0041) < 0.0001


def test_ensure_sound_speed_rule():
    assert O.ensure_sound_speed(20.0, 1.0) == 20.0
    assert O.ensure_sound_speed(20.0, 3.0) == 30.0
    assert O.ensure_sound_speed(20.0, -5.0) == 20.0
    assert O.ensure_sound_speed(20.0, float("nan")) == 20.0


# --- Monaghan viscosity (viscosity.rs:51-73; formula from its module doc, 84-104) ---
def test_pi_ij_published_form_and_properties():
    alpha, beta, eps, h, c = 0.05, 0.3, 0.01, 0.01, 20.0
    xi, xj = [0.001, 0.002, 0.003], [0.0085, -0.001, 0.004]
    vi, vj = [0.5, -0.1, 0.2], [-0.3, 0.1, 0.0]
    rho_i, rho_j = 1003.0, 998.5
    got = O.pi_ij(alpha, beta, eps, h, c, xi, xj, vi, vj, rho_i, rho_j)
    # mu = h (v_ij . r_ij)/(r^2 + eps h^2), Pi = (-alpha c mu + beta mu^2)/rho_bar (Monaghan 1992)
    fx = [F(a) - F(b) for a, b in zip(xi, xj)]
    fv = [F(a) - F(b) for a, b in zip(vi, vj)]
    vr = sum(a * b for a, b in zip(fv, fx))
    assert vr < 0
    r2 = sum(a * a for a in fx)
    mu = F(h) * vr / (r2 + F(eps) * F(h) * F(h))
    ex = (-F(alpha) * F(c) * mu + F(beta) * mu * mu) / ((F(rho_i) + F(rho_j)) / 2)
    assert abs(got - float(ex)) <= 1e-14 * abs(float(ex))
    assert got > 0
    # symmetric under i <-> j
    assert O.pi_ij(alpha, beta, eps, h, c, xj, xi, vj, vi, rho_j, rho_i) == got
    # separating pair: zero
    assert O.pi_ij(alpha, beta, eps, h, c, xi, xj, vj, vi, rho_i, rho_j) == 0.0


def test_reference_kat_viscosity_does_not_increase_ke():
    # tests/fluid_analytic.rs:171-209
    dx = 0.08
    h = 2.2 * dx
    rho0 = 1000.0
    mass = O.mass_from_spacing(rho0, dx, 3)
    x0, x1 = np.array([0.0, 0, 0]), np.array([dx * 0.8, 0, 0])
    v0, v1 = np.array([1.0, 0, 0]), np.array([-1.0, 0, 0])
    ke0 = 0.5 * mass * (v0 @ v0) + 0.5 * mass * (v1 @ v1)
    gw = O.grad_w(x1 - x0, h, 3)
    pi = O.pi_ij(0.5, 0.0, 0.01, h, 40.0, x0, x1, v0, v1, rho0, rho0)
    f = gw * (-mass * pi)
    dt = 1e-4
    v0n, v1n = v0 + f * dt, v1 - f * dt
    ke1 = 0.5 * mass * (v0n @ v0n) + 0.5 * mass * (v1n @ v1n)
    assert ke1 <= ke0 + 1e-9


# --- floor_i32 (grid.rs:144-157) ------------------------------------------------------
def test_floor_i32_edge_cases():
    cases = {0.0: 0, -0.0: 0, 0.5: 0, -0.5: -1, -1.0: -1, 2.9999: 2, -2.0000001: -3,
             1e300: 2147483646, -1e300: -2147483647, float("nan"): 0, float("inf"): 0, float("-inf"): 0,
             2147483646.5: 2147483646, -2147483646.5: -2147483647}
    for x, k in cases.items():
        assert O.floor_i32(x) == k, (x, O.floor_i32(x), k)


# --- published vectors ---------------------------------------------------------------
def test_fnv1a64_published_vectors():
    # FNV-1a 64 (Fowler/Noll/Vo), IETF draft-eastlake-fnv test vectors.
    assert O.fnv1a_bytes(b"") == 0xCBF29CE484222325
    assert O.fnv1a_bytes(b"a") == 0xAF63DC4C8601EC8C
    assert O.fnv1a_bytes(b"foobar") == 0x85944171F73967E8
    vals = np.array([1.5, -2.25, 3.0e-7])
    assert O.fnv1a_f64(vals) == O.fnv1a_bytes(vals.astype("<f8").tobytes())


def test_xoshiro256starstar_published_vector():
    # Blackman & Vigna reference implementation from state {1, 2, 3, 4}.
    r = O.Rng(0)
    r.set_state([1, 2, 3, 4])
    want = [11520, 0, 1509978240, 1215971899390074240, 1216172134540287360, 607988272756665600,
            16172922978634559625, 8476171486693032832, 10595114339597558777, 2904607092377533576]
    assert [r.next_u64() for _ in range(10)] == want


def test_splitmix64_published_vector():
    # SplitMix64 (Steele, Lea & Flood) from state 0.
    assert O.splitmix64_stream(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_range():
    r = O.Rng(7)
    xs = [r.uniform(-0.5, 0.25) for _ in range(2000)]
    assert min(xs) >= -0.5 and max(xs) < 0.25
