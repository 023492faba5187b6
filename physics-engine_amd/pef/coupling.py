"""Two-way fluid-rigid coupling around the device WCSPH step (coupling.rs).

`CoupledWcsphWorld.step` keeps the reference's order (coupling.rs:853-865):

1. the coupler places every rigid sample from its body's pose and motion
   (`FluidRigidCoupler::sync_from_rigid`, on the device: pef_world_sync_rigid);
2. the fluid steps (pef_world_step);
3. the samples' boundary reactions are reduced to a force and a torque about
   each body's centre (`reduce_with_state`, on the device, FP64:
   pef_world_reduce_rigid);
4. the wrenches are loaded into the rigid world as external loads for one
   symplectic-Euler step, then cleared.

The rigid world is the caller's (pe-core's `World`): `RigidWorld` below is a
host FP64 restatement of the parts the coupled scenes use (world.rs,
integrators.rs:75-123, body.rs, forces.rs `UniformGravity`,
external_loads.rs), operation for operation, so a Python caller can run the
coupled scenes end to end. A Rust caller keeps pe-core and binds only the
three device calls.

Scene builders and measurements follow coupling.rs:412-934 and the
reference's coupling oracle tests (tests/coupling_oracles.rs).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

STANDARD_GRAVITY = 9.80665  # units.rs:7

Vec = Tuple[float, float, float]
Mat = Tuple[Vec, Vec, Vec]
Quat = Tuple[float, float, float, float]  # (w, x, y, z)


# --- pe-math / pe-core arithmetic, in the reference's operation order ----------
def _add(a: Vec, b: Vec) -> Vec:
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def _sub(a: Vec, b: Vec) -> Vec:
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _mul(a: Vec, s: float) -> Vec:
    return (a[0] * s, a[1] * s, a[2] * s)


def _cross(a: Vec, b: Vec) -> Vec:  # vec3.rs:87-93
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _norm(a: Vec) -> float:
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def _mat_vec(m: Mat, v: Vec) -> Vec:  # body.rs:233-239
    return (m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2],
            m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
            m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2])


def _mat_mul(a: Mat, b: Mat) -> Mat:  # body.rs:243-251
    return tuple(tuple(a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j] for j in range(3))
                 for i in range(3))


def _transpose(m: Mat) -> Mat:
    return tuple(tuple(m[j][i] for j in range(3)) for i in range(3))


def invert3(m: Mat) -> Mat:
    """body.rs:204-229: cofactors times 1/det; singular -> ValueError."""
    det = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
           - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
           + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))
    if not (math.isfinite(det) and abs(det) > 0.0):
        raise ValueError("singular matrix")
    d = 1.0 / det
    return (((m[1][1] * m[2][2] - m[1][2] * m[2][1]) * d, (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * d,
             (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * d),
            ((m[1][2] * m[2][0] - m[1][0] * m[2][2]) * d, (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * d,
             (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * d),
            ((m[1][0] * m[2][1] - m[1][1] * m[2][0]) * d, (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * d,
             (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * d))


def quat_to_matrix(q: Quat) -> Mat:  # quat.rs:119-138
    w, x, y, z = q
    return ((1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y)),
            (2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x)),
            (2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)))


def quat_rotate(q: Quat, v: Vec) -> Vec:  # quat.rs:102-109
    return _mat_vec(quat_to_matrix(q), v)


def quat_multiply(a: Quat, b: Quat) -> Quat:  # quat.rs:81-90
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return (aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
            aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw)


def quat_normalize(q: Quat) -> Quat:  # quat.rs:62-77 (non-finite or zero -> identity)
    if not all(math.isfinite(c) for c in q):
        return (1.0, 0.0, 0.0, 0.0)
    n = math.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3])
    if n <= 0.0:
        return (1.0, 0.0, 0.0, 0.0)
    return (q[0] / n, q[1] / n, q[2] / n, q[3] / n)


def quat_exp_rotvec(rv: Vec) -> Quat:  # quat.rs:152-162
    th = _norm(rv)
    half = 0.5 * th
    s = math.sin(half) / th if th > 1e-8 else 0.5 - th * th / 48.0
    return (math.cos(half), s * rv[0], s * rv[1], s * rv[2])


def inertia_solid_box(mass: float, a: float, b: float, c: float) -> Mat:  # body.rs:181-187
    return ((mass * (b * b + c * c) / 12.0, 0.0, 0.0), (0.0, mass * (a * a + c * c) / 12.0, 0.0),
            (0.0, 0.0, mass * (a * a + b * b) / 12.0))


def inertia_solid_sphere(mass: float, radius: float) -> Mat:  # body.rs:175-178
    i = (2.0 / 5.0) * mass * radius * radius
    return ((i, 0.0, 0.0), (0.0, i, 0.0), (0.0, 0.0, i))


# --- rigid bodies (pe-core) ---------------------------------------------------------
@dataclass
class RigidBody:
    """RigidBody (body.rs:17-61): identity inertia and pose, at rest."""

    mass: float
    inertia_body: Mat = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))
    position: Vec = (0.0, 0.0, 0.0)
    velocity: Vec = (0.0, 0.0, 0.0)
    orientation: Quat = (1.0, 0.0, 0.0, 0.0)
    angular_velocity: Vec = (0.0, 0.0, 0.0)

    def __post_init__(self):
        if not (math.isfinite(self.mass) and self.mass > 0.0):
            raise ValueError("mass must be finite and positive")

    def with_position(self, p) -> "RigidBody":
        self.position = tuple(float(c) for c in p)
        return self

    def with_velocity(self, v) -> "RigidBody":
        self.velocity = tuple(float(c) for c in v)
        return self

    def with_orientation(self, q) -> "RigidBody":
        self.orientation = quat_normalize(tuple(float(c) for c in q))
        return self

    def with_angular_velocity(self, w) -> "RigidBody":
        self.angular_velocity = tuple(float(c) for c in w)
        return self

    def with_inertia(self, i) -> "RigidBody":
        m = tuple(tuple(float(c) for c in row) for row in i)
        if not all(math.isfinite(c) for row in m for c in row):
            raise ValueError("inertia must be finite")
        if any(m[a][b] != m[b][a] for a in range(3) for b in range(3)):
            raise ValueError("inertia must be symmetric")
        if not (np.linalg.eigvalsh(np.array(m)) > 0.0).all():
            raise ValueError("inertia must be positive definite")
        self.inertia_body = m
        return self


class RigidWorld:
    """pe-core World (world.rs:63-215) with the generators the coupled scenes
    install, in their order: UniformGravity (forces.rs:43-48), then
    ExternalLoads (external_loads.rs:150-160). State per body: x, q (w, x, y,
    z), v and the world-frame angular momentum L."""

    def __init__(self, bodies: Sequence[RigidBody], gravity=(0.0, 0.0, 0.0)):
        self.n = len(bodies)
        self.mass = [b.mass for b in bodies]
        self.inertia_body = [b.inertia_body for b in bodies]
        self.inv_inertia_body = [invert3(b.inertia_body) for b in bodies]
        self.gravity: Vec = tuple(float(c) for c in gravity)
        self.x: List[Vec] = []
        self.q: List[Quat] = []
        self.v: List[Vec] = []
        self.l: List[Vec] = []
        for b in bodies:
            qq = quat_normalize(b.orientation)
            r = quat_to_matrix(qq)
            i_w = _mat_mul(_mat_mul(r, b.inertia_body), _transpose(r))
            self.x.append(tuple(b.position))
            self.q.append(qq)
            self.v.append(tuple(b.velocity))
            self.l.append(_mat_vec(i_w, tuple(b.angular_velocity)))
        self.t = 0.0
        self.load_force: List[Vec] = [(0.0, 0.0, 0.0)] * self.n
        self.load_torque: List[Vec] = [(0.0, 0.0, 0.0)] * self.n

    def inv_inertia_world(self, q: Quat, i: int) -> Mat:  # world.rs:112-119
        r = quat_to_matrix(q)
        return _mat_mul(_mat_mul(r, self.inv_inertia_body[i]), _transpose(r))

    def omega_i(self, i: int) -> Vec:  # world.rs:129-131
        return _mat_vec(self.inv_inertia_world(self.q[i], i), self.l[i])

    def omega(self) -> List[Vec]:
        return [self.omega_i(i) for i in range(self.n)]

    def set_loads(self, force, torque):
        """ExternalLoads::set_all (external_loads.rs:100-106)."""
        self.load_force = [tuple(float(c) for c in f) for f in force]
        self.load_torque = [tuple(float(c) for c in t) for t in torque]

    def clear_loads(self):
        self.load_force = [(0.0, 0.0, 0.0)] * self.n
        self.load_torque = [(0.0, 0.0, 0.0)] * self.n

    def forces(self):  # world.rs:134-141
        f = [(0.0, 0.0, 0.0)] * self.n
        tau = [(0.0, 0.0, 0.0)] * self.n
        for i in range(self.n):
            f[i] = _add(f[i], _mul(self.gravity, self.mass[i]))
        for i in range(self.n):
            if i < len(self.load_force):
                f[i] = _add(f[i], self.load_force[i])
            if i < len(self.load_torque):
                tau[i] = _add(tau[i], self.load_torque[i])
        return f, tau

    def step(self, dt: float):
        """World::step (world.rs:208-215) with symplectic_euler
        (integrators.rs:108-123) and drift_rotation's implicit midpoint with
        four fixed-point iterations (integrators.rs:75-100)."""
        f, tau = self.forces()
        v = [_add(self.v[i], _mul(f[i], dt / self.mass[i])) for i in range(self.n)]
        l = [_add(self.l[i], _mul(tau[i], dt)) for i in range(self.n)]
        q = []
        for i in range(self.n):
            qi, li = self.q[i], l[i]
            w = _mat_vec(self.inv_inertia_world(qi, i), li)
            for _ in range(4):
                q_mid = quat_multiply(quat_exp_rotvec(_mul(w, 0.5 * dt)), qi)
                w = _mat_vec(self.inv_inertia_world(q_mid, i), li)
            q.append(quat_normalize(quat_multiply(quat_exp_rotvec(_mul(w, dt)), qi)))
        self.x = [_add(self.x[i], _mul(v[i], dt)) for i in range(self.n)]
        self.v, self.l = v, l
        self.q = [quat_normalize(qq) for qq in q]
        self.t += dt

    def pack(self) -> np.ndarray:
        """State::pack (world.rs:35-50): x..., q..., v..., L..."""
        return np.array([c for a in self.x for c in a] + [c for a in self.q for c in a]
                        + [c for a in self.v for c in a] + [c for a in self.l for c in a], dtype=np.float64)


# --- FluidRigidCoupler (coupling.rs:70-349) ------------------------------------------
@dataclass
class RigidSample:
    """RigidSample (coupling.rs:43-53)."""

    body: int
    local: Vec
    local_n: Vec
    particle: int


def box_samples(dim: int, half, dx: float, layers: int = 1):
    """Body-frame sample positions and outward normals of an axis-aligned box
    surface in FluidRigidCoupler::sample_box's order (coupling.rs:104-230): per
    layer, the y faces then the x faces in 2-D (normals zero: 2-D box faces are
    unfiltered); the z, y, x faces in 3-D."""
    hx, hy, hz = (float(c) for c in half)
    layers = max(int(layers), 1)
    nx = max(_round_half_away(2.0 * hx / dx), 1)
    ny = max(_round_half_away(2.0 * hy / dx), 1)
    xs: List[Vec] = []
    ns: List[Vec] = []

    def push(p, n):
        xs.append(p)
        ns.append(n)

    for layer in range(layers):
        off = layer * dx
        if dim == 2:
            for i in range(nx + 1):
                x = -hx + i * (2.0 * hx) / nx
                push((x, -hy - off, 0.0), (0.0, 0.0, 0.0))
                push((x, hy + off, 0.0), (0.0, 0.0, 0.0))
            for j in range(1, ny):
                y = -hy + j * (2.0 * hy) / ny
                push((-hx - off, y, 0.0), (0.0, 0.0, 0.0))
                push((hx + off, y, 0.0), (0.0, 0.0, 0.0))
            continue
        nz = max(_round_half_away(2.0 * hz / dx), 1)
        for i in range(nx + 1):
            for j in range(ny + 1):
                x = -hx + i * (2.0 * hx) / nx
                y = -hy + j * (2.0 * hy) / ny
                push((x, y, -hz - off), (0.0, 0.0, -1.0))
                push((x, y, hz + off), (0.0, 0.0, 1.0))
        for i in range(nx + 1):
            for k in range(1, nz):
                x = -hx + i * (2.0 * hx) / nx
                z = -hz + k * (2.0 * hz) / nz
                push((x, -hy - off, z), (0.0, -1.0, 0.0))
                push((x, hy + off, z), (0.0, 1.0, 0.0))
        for j in range(1, ny):
            for k in range(1, nz):
                y = -hy + j * (2.0 * hy) / ny
                z = -hz + k * (2.0 * hz) / nz
                push((-hx - off, y, z), (-1.0, 0.0, 0.0))
                push((hx + off, y, z), (1.0, 0.0, 0.0))
    return xs, ns


def sphere_samples(dim: int, radius: float, dx: float):
    """FluidRigidCoupler::sample_sphere (coupling.rs:233-265): a circle of
    max(8, round(2 pi r / dx)) samples in 2-D, a golden spiral of
    max(12, round(4 pi r^2 / dx^2)) in 3-D, radial normals."""
    xs: List[Vec] = []
    ns: List[Vec] = []
    if dim == 2:
        n = max(_round_half_away(2.0 * math.pi * radius / dx), 8)
        for i in range(n):
            th = 2.0 * math.pi * i / n
            xs.append((radius * math.cos(th), radius * math.sin(th), 0.0))
            ns.append((math.cos(th), math.sin(th), 0.0))
        return xs, ns
    area = 4.0 * math.pi * radius * radius
    n = int(max(float(_round_half_away(area / (dx * dx))), 12.0))
    ga = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(n):
        y = 1.0 - 2.0 * (i + 0.5) / n
        r_xy = math.sqrt(max(1.0 - y * y, 0.0))
        th = ga * i
        xs.append((radius * r_xy * math.cos(th), radius * y, radius * r_xy * math.sin(th)))
        ns.append((r_xy * math.cos(th), y, r_xy * math.sin(th)))
    return xs, ns


def _round_half_away(x: float) -> int:
    """f64::round (half away from zero) for the non-negative counts here."""
    return int(math.floor(x + 0.5)) if x >= 0.0 else -int(math.floor(-x + 0.5))


class FluidRigidCoupler:
    """Sample table of the rigid bodies' Akinci boundary particles
    (coupling.rs:57-103). Samples are appended to a host `pef.Particles` at
    their body-frame positions, as the reference appends them; the scene
    builder then places them at the body's pose."""

    RIGID_PSI_SCALE = 1.0  # coupling.rs:98-103

    def __init__(self, n_bodies: int = 0):
        self.samples: List[RigidSample] = []
        self.n_bodies = n_bodies

    def resize_bodies(self, n: int):
        self.n_bodies = n

    def _psi(self, dim: int, dx: float) -> float:
        return (dx * dx if dim == 2 else dx * dx * dx) * self.RIGID_PSI_SCALE

    def _push(self, parts, body: int, xs, ns, psi: float):
        from . import BOUNDARY

        first = len(parts)
        m = len(xs)
        parts._append(np.asarray(xs, float).reshape(-1, 3), np.zeros((m, 3)), np.full(m, BOUNDARY, np.uint8),
                      np.full(m, psi), np.asarray(ns, float).reshape(-1, 3))
        for k in range(m):
            self.samples.append(RigidSample(body, tuple(xs[k]), tuple(ns[k]), first + k))
        self.n_bodies = max(self.n_bodies, body + 1)
        return m

    def sample_box(self, parts, body: int, half, dx: float, layers: int = 1) -> int:
        xs, ns = box_samples(parts.dim, half, dx, layers)
        return self._push(parts, body, xs, ns, self._psi(parts.dim, dx))

    def sample_sphere(self, parts, body: int, radius: float, dx: float) -> int:
        xs, ns = sphere_samples(parts.dim, radius, dx)
        return self._push(parts, body, xs, ns, self._psi(parts.dim, dx))

    def arrays(self):
        """(particle [ns] u64, body [ns] u32, local [ns][3], local_n [ns][3])."""
        s = self.samples
        return (np.array([a.particle for a in s], np.uint64), np.array([a.body for a in s], np.uint32),
                np.array([a.local for a in s], float).reshape(-1, 3),
                np.array([a.local_n for a in s], float).reshape(-1, 3))

    def place(self, parts, rigid: RigidWorld):
        """Put the samples of a scene under construction at their bodies' poses
        (x + R local, velocity v + w x r, 2-D z zeroed), as the scene builders
        do before the fluid world exists (coupling.rs:919-926)."""
        for s in self.samples:
            r = quat_rotate(rigid.q[s.body], s.local)
            w = rigid.omega_i(s.body)
            p = _add(rigid.x[s.body], r)
            v = _add(rigid.v[s.body], _cross(w, r))
            if parts.dim == 2:
                p, v = (p[0], p[1], 0.0), (v[0], v[1], 0.0)
            parts.x[s.particle] = p
            parts.v[s.particle] = v


def carve_fluid_box(parts, centre, half, margin: float):
    """carve_fluid_box (coupling.rs:561-600): drop the fluid particles strictly
    inside the box grown by `margin`; every other particle keeps its order."""
    from . import FLUID

    c = np.asarray(centre, float)
    h = np.asarray(half, float) + margin
    d = np.abs(parts.x - c)
    inside = (d[:, 0] < h[0]) & (d[:, 1] < h[1])
    if parts.dim != 2:
        inside &= d[:, 2] < h[2]
    keep = ~(inside & (parts.kind == FLUID))
    parts.x, parts.v, parts.kind = parts.x[keep], parts.v[keep], parts.kind[keep]
    parts.psi, parts.n_out = parts.psi[keep], parts.n_out[keep]
    return int((~keep).sum())


class CoupledWcsphWorld:
    """CoupledWcsphWorld (coupling.rs:812-866): a device WcsphWorld, the
    coupler's samples registered with it once, and the caller's rigid world."""

    def __init__(self, fluid, bodies: Sequence[RigidBody], coupler: FluidRigidCoupler, gravity):
        coupler.resize_bodies(len(bodies))
        self.fluid = fluid
        self.coupler = coupler
        self.rigid = RigidWorld(bodies, gravity)
        particle, body, local, local_n = coupler.arrays()
        fluid.set_rigid_samples(particle, body, local, local_n, self.rigid.n)
        self.force = np.zeros((self.rigid.n, 3))
        self.torque = np.zeros((self.rigid.n, 3))

    def sync_from_rigid(self):
        """FluidRigidCoupler::sync_from_rigid (coupling.rs:289-305) on the device."""
        R = self.rigid
        self.fluid.sync_rigid(np.array(R.x), np.array(R.q), np.array(R.v), np.array(R.omega()))

    def reduce_with_state(self):
        """FluidRigidCoupler::reduce_with_state (coupling.rs:308-330) on the device."""
        self.force, self.torque = self.fluid.reduce_rigid(np.array(self.rigid.x))
        return self.force, self.torque

    def step(self, dt: float):
        """coupling.rs:853-865."""
        self.sync_from_rigid()
        self.fluid.step(dt)
        f, t = self.reduce_with_state()
        self.rigid.set_loads(f, t)
        self.rigid.step(dt)
        self.rigid.clear_loads()


class CoupledDfsphWorld(CoupledWcsphWorld):
    """CoupledDfsphWorld (coupling.rs:353-408) over a device DfsphWorld. Its
    step is the WCSPH one's: pef_world_sync_rigid also brings the lists,
    densities and alpha of the moved state up to date, which is the
    reference's rebuild_neighbours, compute_density, compute_alpha."""


# --- scenes and measurements (coupling.rs:412-441, 603-981) ----------------------
def archimedes_box_draft(rho_body: float, rho_fluid: float, height: float) -> float:  # coupling.rs:412-415
    return (rho_body / max(rho_fluid, 1e-12)) * height


def heave_period_analytic(mass: float, rho_fluid: float, a_wp: float, g: float) -> float:  # coupling.rs:418-421
    return 2.0 * math.pi * math.sqrt(mass / max(rho_fluid * g * a_wp, 1e-12))


def barge_heave_added_mass_2d(rho_fluid: float, beam: float) -> float:  # coupling.rs:424-428
    return rho_fluid * math.pi * beam * beam / 8.0


def heave_period_with_added_mass(mass: float, m_a: float, rho_fluid: float, a_wp: float,
                                 g: float) -> float:  # coupling.rs:430-434
    return 2.0 * math.pi * math.sqrt((mass + m_a) / max(rho_fluid * g * a_wp, 1e-12))


def sphere_freefall_accel_added_mass(rho_s: float, rho_f: float, g: float) -> float:  # coupling.rs:437-441
    return g * (rho_s - rho_f) / max(rho_s + 0.5 * rho_f, 1e-12)


def estimate_heave_period(ys) -> Optional[float]:
    """Mean spacing of the up-crossings of the demeaned y(t), linearly
    interpolated (coupling.rs:715-740); None with fewer than 8 samples or 2
    crossings."""
    if len(ys) < 8:
        return None
    mean = sum(y for _, y in ys) / len(ys)
    cross = []
    for i in range(1, len(ys)):
        y0, y1 = ys[i - 1][1] - mean, ys[i][1] - mean
        if y0 <= 0.0 and y1 > 0.0:
            t0, t1 = ys[i - 1][0], ys[i][0]
            frac = -y0 / max(y1 - y0, 1e-15)
            cross.append(t0 + frac * (t1 - t0))
    if len(cross) < 2:
        return None
    s = 0.0
    for i in range(1, len(cross)):
        s += cross[i] - cross[i - 1]
    return s / (len(cross) - 1)


def ensemble_heave_period(runs, f_min: float = 0.5, f_max: float = 3.0, pad: int = 8) -> float:
    """Dominant period of an ensemble of heave records. No counterpart in the
    reference, whose estimator (estimate_heave_period) reads the up-crossings of
    one record; at coarse resolution those move with the last bits of the state.
    Each record is a list of (t, y) samples, all records on the same uniform
    times. Each y is detrended by its least-squares line, Hann-windowed and
    zero-padded `pad` times; the power spectra of the records are summed, and
    the period is 1 / (the frequency of the largest summed power in
    [f_min, f_max] Hz)."""
    total, fr, t_ref = None, None, None
    for ys in runs:
        a = np.asarray(ys, dtype=np.float64)
        t, y = a[:, 0] - a[0, 0], a[:, 1]
        if t_ref is None:
            t_ref = t
        elif len(t) != len(t_ref) or not np.allclose(t, t_ref, rtol=0.0, atol=1e-9):
            raise ValueError("heave records must share their sample times")
        yd = y - np.polyval(np.polyfit(t, y, 1), t)
        n = len(yd)
        p = np.abs(np.fft.rfft(yd * np.hanning(n), pad * n)) ** 2
        fr = np.fft.rfftfreq(pad * n, t[1] - t[0])
        total = p if total is None else total + p
    if total is None:
        raise ValueError("no heave records")
    m = (fr >= f_min) & (fr <= f_max)
    return float(1.0 / fr[m][np.argmax(total[m])])


def make_floating_box_wcsph_2d(dx: float, tank_w: float, water_h: float, hx: float, hy: float,
                               rho_body: float, options=None) -> CoupledWcsphWorld:
    """make_floating_box_wcsph_2d (coupling.rs:869-934): a 2-D tank of water
    (two-layer Akinci shell, recomputed psi), a box of density rho_body at its
    Archimedes draft (+dx/4) with the water inside it carved out (margin dx),
    WCSPH with c = 40, alpha = 0.15, XSPH 0.1, no sound-speed adaptation."""
    import pef

    rho0 = 1000.0
    hsm = 2.2 * dx
    g = STANDARD_GRAVITY
    parts = pef.Particles(pef.D2, rho0, pef.mass_calibrated_to_rest(rho0, dx, hsm, pef.D2), hsm)
    nx = max(int(math.floor(tank_w / dx)), 4)
    ny = max(int(math.floor(water_h / dx)), 4)
    pef.fill_box(parts, (0.5 * dx, 0.5 * dx, 0.0), nx, ny, 1, dx)
    pef.fill_box_boundary(parts, (0.0, 0.0, 0.0), (nx * dx, water_h * 2.5, 0.0), dx, 2)
    pef.recomputed_akinci_psi(parts)
    draft = archimedes_box_draft(rho_body, rho0, 2.0 * hy)
    y_com = (water_h - draft) + hy + 0.25 * dx
    x_com = 0.5 * tank_w
    half = (hx, hy, 0.0)
    carve_fluid_box(parts, (x_com, y_com, 0.0), half, 1.0 * dx)
    coupler = FluidRigidCoupler(1)
    coupler.sample_box(parts, 0, half, dx, 1)
    mass_b = rho_body * (4.0 * hx * hy)
    body = RigidBody(mass_b).with_position((x_com, y_com, 0.0)).with_inertia(
        inertia_solid_box(mass_b, 2.0 * hx, 2.0 * hy, 1.0))
    # the builder places the samples at rest (velocity zero) before the world exists
    rigid = RigidWorld([body], (0.0, -g, 0.0))
    for s in coupler.samples:
        r = quat_rotate(rigid.q[0], s.local)
        p = _add(rigid.x[0], r)
        parts.x[s.particle] = (p[0], p[1], 0.0)
        parts.v[s.particle] = (0.0, 0.0, 0.0)
    params = pef.WcsphParams.water(rho0, 40.0, (0.0, -g, 0.0))
    params.adapt_sound_speed = False
    params.xsph_epsilon = 0.1
    params.visc_alpha = 0.15
    fluid = pef.WcsphWorld(parts, params, options)
    w = CoupledWcsphWorld(fluid, [body], coupler, (0.0, -g, 0.0))
    w.particles, w.params = parts, params
    return w


def make_floating_box_2d(dx: float, tank_w: float, water_h: float, hx: float, hy: float, rho_body: float,
                         options=None) -> CoupledDfsphWorld:
    """make_floating_box_2d (coupling.rs:603-686): the DFSPH floating box, the
    body at its Archimedes draft + dx/2, the water inside it carved out
    (margin 1.25 dx), tolerances 1 %, 40 + 25 iterations, XSPH 0.15, no warm
    start, velocity cap 0.8, free-slip planes x >= 0, y >= 0, x <= nx dx."""
    import pef

    rho0 = 1000.0
    hsm = 2.2 * dx
    g = STANDARD_GRAVITY
    parts = pef.Particles(pef.D2, rho0, pef.mass_calibrated_to_rest(rho0, dx, hsm, pef.D2), hsm)
    nx = max(int(math.floor(tank_w / dx)), 4)
    ny = max(int(math.floor(water_h / dx)), 4)
    pef.fill_box(parts, (0.5 * dx, 0.5 * dx, 0.0), nx, ny, 1, dx)
    pef.fill_box_boundary(parts, (0.0, 0.0, 0.0), (nx * dx, water_h * 2.5, 0.0), dx, 2)
    pef.recomputed_akinci_psi(parts)
    draft = archimedes_box_draft(rho_body, rho0, 2.0 * hy)
    y_com = (water_h - draft) + hy + 0.5 * dx
    x_com = 0.5 * tank_w
    half = (hx, hy, 0.0)
    carve_fluid_box(parts, (x_com, y_com, 0.0), half, 1.25 * dx)
    coupler = FluidRigidCoupler(1)
    coupler.sample_box(parts, 0, half, dx, 1)
    mass_b = rho_body * (4.0 * hx * hy)
    body = RigidBody(mass_b).with_position((x_com, y_com, 0.0)).with_inertia(
        inertia_solid_box(mass_b, 2.0 * hx, 2.0 * hy, 1.0))
    rigid = RigidWorld([body], (0.0, -g, 0.0))
    coupler.place(parts, rigid)
    params = pef.DfsphParams(gravity=(0.0, -g, 0.0), particle_diameter=dx, density_error_tol=0.01,
                             divergence_error_tol=0.01, max_density_iters=40, max_divergence_iters=25,
                             xsph_epsilon=0.15, adaptive_dt=False, warm_start=False, max_velocity_correction=0.8,
                             free_slip_xmin=0.0, free_slip_ymin=0.0, free_slip_xmax=nx * dx)
    fluid = pef.DfsphWorld(parts, params, options)
    w = CoupledDfsphWorld(fluid, [body], coupler, (0.0, -g, 0.0))
    w.particles, w.params = parts, params
    return w


def floating_box_draft_error_wcsph(world: CoupledWcsphWorld, hy: float, water_h: float):
    """coupling.rs:936-961: (draft_sim, draft_ref, rel) with the simulated
    draft from the keel height and the reference draft m / (rho0 A_wp),
    A_wp = 2 max |local x| over body 0's samples."""
    rho0 = world.particles.rest_density
    mass = world.rigid.mass[0]
    keel = world.rigid.x[0][1] - hy
    draft_sim = water_h - keel
    hx = 0.0
    for s in world.coupler.samples:
        if s.body == 0 and abs(s.local[0]) > hx:
            hx = abs(s.local[0])
    draft_ref = mass / max(rho0 * (2.0 * hx), 1e-12)
    rel = abs(draft_sim - draft_ref) / max(draft_ref, 1e-12)
    return draft_sim, draft_ref, rel


def max_box_penetration_wcsph(world: CoupledWcsphWorld, half) -> float:
    """max_box_penetration_wcsph (coupling.rs:963-981): the deepest fluid
    particle inside body 0's box (body frame, x and y as the reference measures
    it; z too in 3-D), 0 when none is inside."""
    from . import FLUID

    x = world.fluid.positions()[world.particles.kind == FLUID]
    q = world.rigid.q[0]
    r = quat_to_matrix((q[0], -q[1], -q[2], -q[3]))  # inverse_rotate = conjugate's rotation
    d = x - np.asarray(world.rigid.x[0])
    loc = np.stack([r[a][0] * d[:, 0] + r[a][1] * d[:, 1] + r[a][2] * d[:, 2] for a in range(3)], 1)
    hx, hy, hz = (float(c) for c in half)
    dx_, dy_ = np.abs(loc[:, 0]) - hx, np.abs(loc[:, 1]) - hy
    if world.particles.dim == 2:
        outside = np.maximum(dx_, dy_)
        depth = np.minimum(-dx_, -dy_)
    else:
        dz_ = np.abs(loc[:, 2]) - hz
        outside = np.maximum(np.maximum(dx_, dy_), dz_)
        depth = np.minimum(np.minimum(-dx_, -dy_), -dz_)
    inside = outside < 0.0
    return float(max(0.0, depth[inside].max())) if inside.any() else 0.0


def coupled_linear_momentum(world: CoupledWcsphWorld) -> np.ndarray:
    """Fluid momentum (diagnostics.rs:39-47) plus the bodies' m v (coupling.rs:802-809)."""
    p = np.array(world.fluid.diagnostics()["momentum"], float)
    for i in range(world.rigid.n):
        p = p + np.asarray(world.rigid.v[i]) * world.rigid.mass[i]
    return p
