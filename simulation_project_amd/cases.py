"""Synthetic input cases for the MPS timestep: dam break, rotating drum, small boxes.

Everything a reference input deck carries (parameter.dat, particle.dat, STL models and the
per-model columns of the configuration file) is produced here in memory, with the reference's
own conventions:
  * particle ids are row indices (Source/IO.cpp:61);
  * triangle normals are recomputed from the winding with the sqrt14 square root
    (Source/IO.cpp:435-453);
  * the rotation angle per step is RPM * dt * 6 degrees (Source/MultiGPUHandler.cpp:690);
  * lambda comes from the reference's piecewise fit (MultiGPUHandler.cpp:1132-1296) and the
    compute zone from the particle + triangle bounding box (MultiGPUHandler.cpp:111-150), both
    through the library's host helpers.
Lattice sites sit a whole number of spacings from the planar walls: the wall weight tables of
the reference (Utility_Inline.h:199-686) count the missing neighbours as lattice layers at
distances r, r + l0, ... behind a wall at distance r, so a first fluid layer at l0 from the wall
starts at the lattice PND n0.  Every site gets a uniform jitter of +-1e-4 l0 per axis from a
counter-based SplitMix64 stream keyed by (seed, particle, axis).
"""
import ctypes as C
import os
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

DEFAULT_SEED = 20261202
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x):
    x = (x + _GOLDEN).astype(np.uint64)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
    return z ^ (z >> np.uint64(31))


def uniform01(seed, index, axis):
    """U[0,1) for each (seed, index, axis); index is an integer array."""
    with np.errstate(over="ignore"):
        ctr = (np.asarray(index, np.uint64) * np.uint64(3) + np.uint64(axis) + np.uint64(1)) * _GOLDEN
        x = splitmix64(np.uint64(seed) ^ ctr)
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def jitter(seed, n, amplitude):
    idx = np.arange(n, dtype=np.uint64)
    return np.stack([(2.0 * uniform01(seed, idx, a) - 1.0) * amplitude for a in range(3)], axis=1)


def sqrt14_f32(x):
    """The reference's approximate square root (solver/Common/inc/Utility_Inline.h:13-26)."""
    x = np.asarray(x, np.float32)
    xhalf = np.float32(0.5) * x
    i = x.view(np.int32)
    y = (np.int32(0x5F3759DF) - (i >> np.int32(1))).astype(np.int32).view(np.float32)
    return (x * y) * (np.float32(1.5) - (xhalf * y) * y)


def triangle_normals(v):
    """CalcTriangleNormal (Source/IO.cpp:435-453) in float32: cross product of (v1-v0, v2-v0)
    divided by its sqrt14 length.  v: float32 [T,3,3]."""
    v = np.asarray(v, np.float32)
    a = v[:, 1] - v[:, 0]
    b = v[:, 2] - v[:, 0]
    nx = a[:, 1] * b[:, 2] - b[:, 1] * a[:, 2]
    ny = a[:, 2] * b[:, 0] - b[:, 2] * a[:, 0]
    nz = a[:, 0] * b[:, 1] - b[:, 0] * a[:, 1]
    ab = sqrt14_f32(nx * nx + ny * ny + nz * nz)
    out = np.zeros((len(v), 3), np.float32)
    ok = ab != np.float32(0.0)
    out[ok, 0] = nx[ok] / ab[ok]
    out[ok, 1] = ny[ok] / ab[ok]
    out[ok, 2] = nz[ok] / ab[ok]
    return out


def default_params(l0, dt, *, grad=2.1, lap=3.1, rho=1000.0, nu=1e-6, g=(0.0, 0.0, -9.81), n0=14.418575,
                   beta=0.97, sonic=20.0, courant=0.2, wall_friction=1.0, explicit_pressure=False,
                   artificial=0.0, r_coeff=0.0, tolerance=0.0, steps=100):
    p = abi.Params()
    p.Dimension = 3
    p.GravityConstant = abi.Float3(*g)
    p.Density = rho
    p.ConstantParticleDensity = n0
    p.InitialParticleDistance = l0
    p.FreeSurfaceThreshold = beta
    p.GradientInfluenceRadiusCoefficient = grad
    p.LaplacianInfluenceRadiusCoefficient = lap
    p.TemperatureInfluenceRadiusCoefficient = grad
    p.FruidSpecificHeat = 1.0
    p.FruidDensity = rho
    p.Dt = dt
    p.AnalysisStep = steps
    p.ThermalStep = 1
    p.NegativeCalculationStep = 1
    p.OutputStep = steps
    p.CourantNumber = courant
    p.WallFrictionCoeffieicnt = wall_friction
    p.StopWallStep = 2 ** 31 - 1
    p.Tolerance = tolerance
    p.bExplicitPressure = 1 if explicit_pressure else 0
    p.SonicVelocity = sonic
    p.ArtificialPressure = artificial
    p.r_coeff = r_coeff
    p.PressureMax = 3.4e38
    p.RLimitCoefficient = 3.4e38
    p.DragPermiabilityConstant = 1.0
    p.MaxNeighborhoodNum = 729
    return p


def tri_params(velocity=(0, 0, 0), rotating=False, rpm=0, dt=0.0, center=(0, 0, 0), second=(0, 0, 0),
               start=0.0, reach=0.0, friction=0.0, is_drag=False, drag_id=-1, temperature=0.0):
    m = abi.TriParams()
    m.Velocity = abi.Float3(*velocity)
    m.Temperature = temperature
    m.RotatingOrNot = 1 if rotating else 0
    # MultiGPUHandler.cpp:690: RotationInDegree = RotationInRPM * Dt * 6 (int * float * int)
    m.AngleOfRotationInDegree = float(np.float32(np.float32(rpm) * np.float32(dt)) * np.float32(6))
    m.RotationStartTime = start
    m.TimeToReachRPM = reach
    m.CenterOfRotation = abi.Float3(*center)
    m.SecondPointOfRotation = abi.Float3(*second)
    m.WallFrictionCoefficient = friction
    m.isDrag = 1 if is_drag else 0
    m.DragTriangleID = drag_id
    return m


def box_triangles(lo, hi):
    """12 triangles of an axis-aligned box (two per face), float32 [12,3,3]."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    c = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float64)
    faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (1, 2, 6, 5), (0, 4, 7, 3)]
    tris = []
    for a, b, cc, d in faces:
        tris.append([c[a], c[b], c[cc]])
        tris.append([c[a], c[cc], c[d]])
    return np.asarray(tris, np.float32)


def _zip_rings(inner, outer):
    """Triangles between two closed rings of points (angles ascending from 0)."""
    ma, mb = len(inner), len(outer)
    tris = []
    i = j = 0
    while i < ma or j < mb:
        a_next = 2.0 * math.pi * (i + 1) / ma
        b_next = 2.0 * math.pi * (j + 1) / mb
        if j == mb or (i < ma and a_next < b_next):
            tris.append([inner[i % ma], outer[j % mb], inner[(i + 1) % ma]])
            i += 1
        else:
            tris.append([inner[i % ma], outer[j % mb], outer[(j + 1) % mb]])
            j += 1
    return tris


def cylinder_triangles(radius, length, center_yz, n_r=16, n_x=36):
    """Closed cylinder along x from 0 to length with near-uniform triangles: each end cap is a
    hexagonal disk (ring j of n_r carries 6 j vertices, 6 n_r^2 triangles per cap) and the mantle
    has 6 n_r segments around and n_x along.  Defaults: 2*1536 + 6912 = 9984 triangles with
    edges of about R / n_r (9.4 mm for R = 0.15 m)."""
    yc, zc = center_yz
    n_theta = 6 * n_r

    def ring(x, r, m):
        th = 2.0 * math.pi * np.arange(m) / m
        return [(x, yc + r * math.cos(t), zc + r * math.sin(t)) for t in th]
    tris = []
    xs = np.linspace(0.0, length, n_x + 1)
    for i in range(n_x):
        a, b = ring(xs[i], radius, n_theta), ring(xs[i + 1], radius, n_theta)
        for k in range(n_theta):
            k2 = (k + 1) % n_theta
            tris.append([a[k], b[k], b[k2]])
            tris.append([a[k], b[k2], a[k2]])
    for x in (0.0, length):
        centre = (x, yc, zc)
        r1 = ring(x, radius / n_r, 6)
        for k in range(6):
            tris.append([centre, r1[k], r1[(k + 1) % 6]])
        for j in range(2, n_r + 1):
            tris.extend(_zip_rings(ring(x, radius * (j - 1) / n_r, 6 * (j - 1)), ring(x, radius * j / n_r, 6 * j)))
    return np.asarray(tris, np.float32)


@dataclass
class Case:
    name: str
    params: abi.Params
    particles: abi.ParticleArrays
    vertices: np.ndarray             # float32 [T,3,3]
    model_of_triangle: np.ndarray    # int32 [T]
    models: list
    drag: list = field(default_factory=list)
    zone: abi.Box = None
    steps: int = 100
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.particles.n

    def triangles(self):
        T = len(self.vertices)
        arr = (abi.Triangle * max(1, T))()
        nrm = triangle_normals(self.vertices) if T else np.zeros((0, 3), np.float32)
        for t in range(T):
            arr[t].ModelIDNumber = int(self.model_of_triangle[t])
            arr[t].Normal = abi.Float3(*map(float, nrm[t]))
            for k in range(3):
                arr[t].Vertex[k] = abi.Float3(*map(float, self.vertices[t, k]))
        return arr


@dataclass
class Spheres:
    """Rigid spheres carried by the fluid through Schiller-Naumann drag."""
    position: np.ndarray        # [n, 3]
    velocity: np.ndarray        # [n, 3]
    diameter: float
    density: float
    fluid_viscosity: float = 1e-6


def drum_spheres(case, count=2000, diameter_l0=5.0, density=2500.0, seed=DEFAULT_SEED):
    """`count` non-overlapping spheres of diameter diameter_l0 * l0 placed at random (SplitMix64
    stream of `seed`) inside the fluid of a drum case: centres at least d/2 + l0 from the mantle
    and the caps and d/2 below the free surface, centre distances at least d.  Fewer are returned
    when the fluid cannot hold `count` of them after 200 draws per sphere."""
    meta = case.meta
    l0 = meta["l0"]
    d = diameter_l0 * l0
    R, L = meta["radius"], meta["length"]
    yc = zc = R
    h = meta["level"]
    u = (_splitmix_uniform(seed ^ 0x5EED5, count * 200 * 3)).reshape(count * 200, 3)
    cell = d
    grid = {}
    out = []
    k = 0
    for draw in u:
        if len(out) >= count:
            break
        x = d / 2 + l0 + draw[0] * (L - d - 2 * l0)
        ang = 2 * math.pi * draw[1]
        rad = (R - d / 2 - l0) * math.sqrt(draw[2])
        y, z = yc + rad * math.cos(ang), zc + rad * math.sin(ang)
        if z - zc > h - d / 2:
            continue
        key = (int(x // cell), int(y // cell), int(z // cell))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                        if (q[0] - x) ** 2 + (q[1] - y) ** 2 + (q[2] - z) ** 2 < d * d:
                            ok = False
        if ok:
            grid.setdefault(key, []).append((x, y, z))
            out.append((x, y, z))
        k += 1
    pos = np.array(out, np.float64).reshape(-1, 3).astype(np.float32).astype(np.float64)
    return Spheres(pos, np.zeros_like(pos), float(np.float32(d)), density, 1e-6)


def _splitmix_uniform(seed, n):
    """n uniform doubles in [0, 1) from a SplitMix64 stream."""
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def _finish(case, lib):
    """Lambda fit and automatic compute zone through the library's host helpers."""
    case.params.LambdaValue = lib.mpsr_calculate_lambda(C.c_float(case.params.InitialParticleDistance))
    tris = case.triangles()
    zone = abi.Box()
    ps = case.particles.struct()
    st = lib.mpsr_auto_compute_zone(C.byref(case.params), C.byref(ps), tris, len(case.vertices), C.byref(zone))
    if st != 0:
        raise RuntimeError("mpsr_auto_compute_zone failed")
    case.zone = zone
    return case


def _particles(pos, nu):
    n = len(pos)
    pa = abi.ParticleArrays(n)
    pa.id[:] = np.arange(n, dtype=np.int32)
    pa.position[:] = pos.astype(np.float32).astype(np.float64)
    pa.kinetic_viscosity[:] = np.float64(np.float32(nu))
    return pa


def dam_break(lib, l0=0.01, seed=DEFAULT_SEED, explicit_pressure=True, steps=100):
    """Dam break: water column 0.2 x 0.1 x 0.2 m in the corner of a 0.5 x 0.2 x 0.4 m box tank
    given as 12 triangles; l0 = 0.01 m gives a 20 x 10 x 20 lattice (4000 particles) with sites at
    l0, 2 l0, ... from the corner walls."""
    nx, ny, nz = int(round(0.2 / l0)), int(round(0.1 / l0)), int(round(0.2 / l0))
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pos = (np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1) + 1.0) * l0
    pos = pos + jitter(seed, len(pos), 1e-4 * l0)
    p = default_params(l0, 1e-4, explicit_pressure=explicit_pressure,
                       artificial=0.01 if explicit_pressure else 0.0,
                       r_coeff=0.7 if explicit_pressure else 0.0, steps=steps)
    verts = box_triangles((0.0, 0.0, 0.0), (0.5, 0.2, 0.4))
    case = Case("dam_break", p, _particles(pos, 1e-6), verts, np.zeros(len(verts), np.int32),
                [tri_params(friction=1.0)], steps=steps, meta={"l0": l0, "seed": seed})
    return _finish(case, lib)


def _fill_level(fill):
    """Height of the fill line above the axis, over the radius, whose circular segment below holds
    `fill` of the disc (bisection)."""
    lo_h, hi_h = -1.0, 1.0
    for _ in range(200):
        h = 0.5 * (lo_h + hi_h)
        d = -h
        seg = math.acos(max(-1.0, min(1.0, d))) - d * math.sqrt(max(0.0, 1.0 - d * d))
        if seg / math.pi < fill:
            lo_h = h
        else:
            hi_h = h
    return 0.5 * (lo_h + hi_h)


def _drum_lattice(l0, fill, radius, length, count_only=False):
    """Lattice sites of a drum of the given radius and length (axis along x at y = z = radius)
    filled to `fill` of its cross-section: planes at l0, 2 l0, ... and at least l0 from the far
    cap, sites at least l0 inside the mantle.  Returns (positions, fill level h above the axis),
    or only the site count when count_only."""
    yc = zc = radius
    h = _fill_level(fill) * radius
    nx = int(math.floor(length / l0 + 1e-9)) - 1
    m = int(radius / l0)
    off = np.arange(-m, m + 1) * l0
    yy, zz = np.meshgrid(yc + off, zc + off, indexing="ij")
    rr = (yy - yc) ** 2 + (zz - zc) ** 2
    inside = (rr <= (radius - l0) ** 2) & ((zz - zc) < h)
    if count_only:
        return max(nx, 0) * int(inside.sum())
    yz = np.stack([yy[inside], zz[inside]], axis=1)
    xs = (np.arange(nx) + 1.0) * l0
    pos = np.empty((max(nx, 0) * len(yz), 3))
    pos[:, 0] = np.repeat(xs, len(yz))
    pos[:, 1] = np.tile(yz[:, 0], nx)
    pos[:, 2] = np.tile(yz[:, 1], nx)
    return pos, h


def drum(lib, l0=2.5e-3, fill=0.4, rpm=30, radius=0.15, length=0.3, seed=DEFAULT_SEED, dt=5e-5,
         n_r=16, n_x=36, steps=200, explicit_pressure=False):
    """Rotating drum: cylinder of radius 0.15 m and length 0.3 m along x (9984 triangles by
    default) spinning at rpm about its axis, filled to the given area fraction.  Lattice sites
    are l0 apart, at l0, 2 l0, ... from the end caps (the last plane at least l0 from the far
    cap) and at least l0 inside the mantle.  l0 = 2.5e-3 m gives 516,936 particles and
    l0 = 2.012e-3 m gives 1,008,472 (the 1M size)."""
    yc = zc = radius
    pos, h = _drum_lattice(l0, fill, radius, length)
    pos = pos + jitter(seed, len(pos), 1e-4 * l0)
    p = default_params(l0, dt, steps=steps, explicit_pressure=explicit_pressure,
                       artificial=0.01 if explicit_pressure else 0.0,
                       r_coeff=0.7 if explicit_pressure else 0.0)
    verts = cylinder_triangles(radius, length, (yc, zc), n_r, n_x)
    m = tri_params(rotating=True, rpm=rpm, dt=dt, center=(0.0, yc, zc), second=(length, yc, zc),
                   friction=1.0)
    case = Case("drum", p, _particles(pos, 1e-6), verts, np.zeros(len(verts), np.int32), [m],
                steps=steps, meta={"l0": l0, "fill": fill, "rpm": rpm, "seed": seed, "radius": radius,
                                   "length": length, "level": h})
    return _finish(case, lib)


def small_box(lib, nx=8, ny=6, nz=8, l0=0.01, seed=DEFAULT_SEED, rotating=False, rpm=60, dt=1e-4,
              explicit_pressure=True, velocity_jitter=0.0):
    """A small water block in a closed box tank, for fast parity tests.  With rotating=True the
    tank spins about the x axis through its centre."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pos = (np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1) + 1.0) * l0
    pos = pos + jitter(seed, len(pos), 1e-4 * l0)
    lo = (0.0, 0.0, 0.0)
    hi = ((nx + 6) * l0, (ny + 4) * l0, (nz + 6) * l0)
    p = default_params(l0, dt, explicit_pressure=explicit_pressure,
                       artificial=0.01 if explicit_pressure else 0.0,
                       r_coeff=0.7 if explicit_pressure else 0.0)
    verts = box_triangles(lo, hi)
    cy, cz = 0.5 * hi[1], 0.5 * hi[2]
    m = tri_params(rotating=rotating, rpm=rpm, dt=dt, center=(0.0, cy, cz), second=(hi[0], cy, cz),
                   friction=1.0)
    pa = _particles(pos, 1e-6)
    if velocity_jitter:
        pa.velocity[:] = (jitter(seed + 1, len(pos), velocity_jitter).astype(np.float32)).astype(np.float64)
    case = Case("small_box", p, pa, verts, np.zeros(len(verts), np.int32), [m], steps=20,
                meta={"l0": l0, "seed": seed})
    return _finish(case, lib)


def moving_plate_box(lib, plate_velocity=(-0.3, 0.0, 0.1), tank_velocity=(0.0, 0.0, 0.05), nx=8, ny=6, nz=8,
                     l0=0.01, seed=DEFAULT_SEED, dt=1e-4, explicit_pressure=True):
    """A small box whose walls translate (UpdateTrianglePosition_Kernel, DeviceProcess_Kernel.h:
    463-482): a vertical paddle (a second, non-drag STL model of two triangles at x = (nx + 2) l0,
    two spacings from the water block's last plane, spanning the tank's y and z) moves with
    plate_velocity toward the water, and the tank itself with tank_velocity.  The paddle's velocity is also the wall velocity of the Laplacian's
    wall term (DeviceProcess_Kernel.h:384) for the particles whose nearest wall it is."""
    case = small_box(lib, nx=nx, ny=ny, nz=nz, l0=l0, seed=seed, dt=dt, explicit_pressure=explicit_pressure)
    xp = (nx + 2) * l0
    y1, z1 = (ny + 4) * l0, (nz + 6) * l0
    plate = np.asarray([[[xp, 0.0, 0.0], [xp, y1, 0.0], [xp, y1, z1]],
                        [[xp, 0.0, 0.0], [xp, y1, z1], [xp, 0.0, z1]]], np.float32)
    case.vertices = np.concatenate([case.vertices, plate])
    case.model_of_triangle = np.concatenate([case.model_of_triangle, np.array([1, 1], np.int32)])
    tank = case.models[0]
    tank.Velocity = abi.Float3(*tank_velocity)
    tank.LinearMovableOrNot = 1 if any(tank_velocity) else 0
    paddle = tri_params(velocity=plate_velocity, friction=1.0)
    paddle.LinearMovableOrNot = 1
    case.models = [tank, paddle]
    case.name = "moving_plate_box"
    return _finish(case, lib)


def create_info(case, precision=abi.MPSR_FP32, pressure_mode=abi.MPSR_PRESSURE_EOS, ppe=None,
                max_particles=None, device=0):
    """Build an mpsr_create_info; returns (info, keepalive)."""
    tris = case.triangles()
    models = (abi.TriParams * max(1, len(case.models)))(*case.models)
    drag = (abi.DragParams * max(1, len(case.drag)))(*case.drag) if case.drag else None
    ps = case.particles.struct()
    info = abi.CreateInfo()
    info.params = case.params
    info.compute_zone = case.zone
    info.max_particles = max_particles or case.n
    info.particles = C.pointer(ps)
    info.triangles = C.cast(tris, C.POINTER(abi.Triangle))
    info.num_triangles = len(case.vertices)
    info.models = C.cast(models, C.POINTER(abi.TriParams))
    info.num_models = len(case.models)
    info.drag = C.cast(drag, C.POINTER(abi.DragParams)) if drag is not None else None
    info.num_drag = len(case.drag)
    info.precision = precision
    info.pressure_mode = pressure_mode
    if ppe is None:
        ppe = abi.PpeOptions(abi.MPSR_PCG_FP64, 1e-8, 20000, abi.MPSR_WARM_ZERO, 0.0, 0.0)
    info.ppe = ppe
    info.device = device
    return info, (tris, models, drag, ps, case)


def bucket_structs(buckets):
    """ctypes array of mpsr_particles for [(ParticleArrays, interval, total), ...] and the objects
    that keep its buffers alive."""
    structs = [b[0].struct() for b in buckets]
    arr = (abi.ParticlesC * max(1, len(structs)))(*structs)
    return arr, (structs, [b[0] for b in buckets])


def drag_box(lib, token="E", nx=8, ny=6, nz=8, l0=0.01, seed=DEFAULT_SEED, dt=1e-4):
    """A small box with a drag plate (an extra two-triangle STL model marked as drag, the
    reference's 'DE V V' / 'DC V' / 'DF V' configuration lines) crossing the water block.  The
    block starts moving with a uniform velocity so the drag terms have a speed to act on."""
    case = small_box(lib, nx=nx, ny=ny, nz=nz, l0=l0, seed=seed, dt=dt)
    x0, x1 = 0.0, (nx + 6) * l0
    zp = (0.5 * nz + 3.3) * l0
    plate = np.asarray([[[x0, 0.0, zp], [x1, 0.0, zp], [x1, (ny + 4) * l0, zp]],
                        [[x0, 0.0, zp], [x1, (ny + 4) * l0, zp], [x0, (ny + 4) * l0, zp]]], np.float32)
    case.vertices = np.concatenate([case.vertices, plate])
    case.model_of_triangle = np.concatenate([case.model_of_triangle, np.array([1, 1], np.int32)])
    case.models = [case.models[0], tri_params(is_drag=True, drag_id=0)]
    d = abi.DragParams()
    d.DragID = 0
    d.DragToken = token.encode()
    d.InputVelocityToken1 = b"V"
    d.OutputVelocityToken2 = b"V"
    d.CoefficientofInflueneRegion = 1.5
    if token == "E":
        d.VelocityIndex = 0.0007
        d.VelocityExponential = 1.5061
        d.VelocityDragPermiabilityConstant = -0.75
    elif token == "B":
        d.DragToken = b"E"
        d.InputVelocityToken1 = b"B"
        d.OutputVelocityToken2 = b"B"
        d.VelocityIndex = 0.0007
        d.VelocityExponential = 1.5061
        d.VelocityDragPermiabilityConstant = -0.75
        d.TempIndex = 0.0005
        d.TempExponential = 0.6668
        d.TempDragPermiabilityConstant = -55.0
    elif token == "C":
        d.ConstantVx, d.ConstantVy, d.ConstantVz = 0.01, 0.0, -0.02
    elif token == "F":
        d.VelocityMagnitueFactor = 0.9
    case.drag = [d]
    case.particles.velocity[:] = np.array([0.05, -0.02, 0.03], np.float32).astype(np.float64)
    case.name = f"drag_box_{token}"
    return _finish(case, lib)


def from_deck(lib, config_path, steps=None):
    """Load one of the reference's input decks (configuration file, parameter file, particle and
    bucket files, STL models) through mpsr_deck_load and return it as a Case.  The deck's own
    set-up is kept as loaded: parameters from the parameter file, lambda from l0, model offsets
    added to the vertices, the compute zone of the configuration or the automatic one.  The
    capacity the reference reserves for bucket inflow is in case.meta["max_particles"], the buckets
    as [(ParticleArrays, interval, total), ...] in case.meta["buckets"]."""
    h = C.c_void_p()
    st = lib.mpsr_deck_load(os.fsencode(config_path), C.byref(h))
    if st != 0:
        raise RuntimeError(f"mpsr_deck_load({config_path}): status {st}: "
                           f"{lib.mpsr_deck_last_error().decode(errors='replace')}")
    try:
        info = abi.CreateInfo()
        if lib.mpsr_deck_create_info(h, C.byref(info)) != 0:
            raise RuntimeError(lib.mpsr_deck_last_error().decode(errors="replace"))
        src = info.particles.contents
        n = src.n
        pa = abi.ParticleArrays(n)
        if n:
            pa.id[:] = np.ctypeslib.as_array(src.id, (n,))
            pa.position[:] = np.ctypeslib.as_array(src.position, (n * 3,)).reshape(n, 3)
            pa.velocity[:] = np.ctypeslib.as_array(src.velocity, (n * 3,)).reshape(n, 3)
            for k in ("pressure", "density", "temperature", "kinetic_viscosity", "solid_phase_rate"):
                getattr(pa, k)[:] = np.ctypeslib.as_array(getattr(src, k), (n,))
            pa.type[:] = np.ctypeslib.as_array(src.type, (n,))
        T = info.num_triangles
        verts = np.zeros((T, 3, 3), np.float32)
        model = np.zeros(T, np.int32)
        for t in range(T):
            tr = info.triangles[t]
            model[t] = tr.ModelIDNumber
            for k in range(3):
                verts[t, k] = (tr.Vertex[k].x, tr.Vertex[k].y, tr.Vertex[k].z)
        models = [abi.TriParams.from_buffer_copy(info.models[m]) for m in range(info.num_models)]
        drag = [abi.DragParams.from_buffer_copy(info.drag[m]) for m in range(info.num_drag)]
        params = abi.Params.from_buffer_copy(info.params)
        zone = abi.Box.from_buffer_copy(info.compute_zone)
        buckets = []
        for j in range(lib.mpsr_deck_num_buckets(h)):
            cnt, itv, tot = C.c_int(), C.c_int(), C.c_int()
            lib.mpsr_deck_bucket(h, j, C.byref(cnt), C.byref(itv), C.byref(tot))
            bp = abi.ParticleArrays(cnt.value)
            bs = bp.struct()
            if lib.mpsr_deck_get_bucket(h, j, C.byref(bs)) != 0:
                raise RuntimeError(lib.mpsr_deck_last_error().decode(errors="replace"))
            buckets.append((bp, itv.value, tot.value))
        meta = {"max_particles": info.max_particles, "buckets": buckets,
                "start_step": lib.mpsr_deck_start_step(h),
                "output_interval": lib.mpsr_deck_output_interval(h),
                "output_ascii": bool(lib.mpsr_deck_output_ascii(h)),
                "output_name": lib.mpsr_deck_output_name(h).decode(errors="replace")}
    finally:
        lib.mpsr_deck_free(h)
    name = os.path.splitext(os.path.basename(config_path))[0]
    return Case(name, params, pa, verts, model, models, drag, zone,
                steps if steps is not None else params.AnalysisStep, meta)


def buckets(lib, count=64, seed=DEFAULT_SEED, l0=5e-3, dt=5e-5, n_min=5e3, n_max=2e5, box_share=0.25):
    """The batched multi-bucket workload: `count` independent rotating buckets, particle counts
    drawn log-uniformly from [n_min, n_max] (64 buckets: about 3.4M particles in all), rotation
    rates from 10-60 rpm, a quarter of them closed boxes (12 triangles), the rest drums of length
    2 R whose radius gives the drawn count at 40 % fill, meshed with up to 9984 triangles.  Every
    bucket uses the reference's explicit (EOS) pressure; l0 = 5e-3 m and dt = 5e-5 s keep the
    sound-speed Courant number c dt / l0 at 0.2.  Deterministic in seed."""
    out = []
    idx = np.arange(count)
    u_n, u_rpm, u_kind = uniform01(seed, idx, 0), uniform01(seed, idx, 1), uniform01(seed, idx, 2)
    for i in range(count):
        target = float(np.exp(np.log(n_min) + u_n[i] * (np.log(n_max) - np.log(n_min))))
        rpm = 10.0 + 50.0 * float(u_rpm[i])
        if u_kind[i] < box_share:
            nx = max(2, int(round((target / 0.75) ** (1.0 / 3.0))))
            ny = max(2, int(round(0.75 * nx)))
            c = small_box(lib, nx=nx, ny=ny, nz=nx, l0=l0, seed=seed + i, rotating=True, rpm=rpm, dt=dt)
        else:
            lo, hi = 4.0 * l0, 2.0
            for _ in range(40):
                r = 0.5 * (lo + hi)
                if _drum_lattice(l0, 0.4, r, 2.0 * r, count_only=True) < target:
                    lo = r
                else:
                    hi = r
            r = 0.5 * (lo + hi)
            n_r = int(min(16, max(2, round(16 * r / 0.15))))
            n_x = int(min(36, max(2, round(36 * 2.0 * r / 0.3))))
            c = drum(lib, l0=l0, fill=0.4, rpm=rpm, radius=r, length=2.0 * r, seed=seed + i, dt=dt, n_r=n_r,
                     n_x=n_x, explicit_pressure=True)
        c.meta.update({"bucket": i, "target": target, "rpm": rpm})
        out.append(c)
    return out


def with_thermal(case, *, temperature=None, wall_temperature=None, conductivity=0.6, resistance=1e-6,
                 wall_conductivity=50.0, wall_resistivity=1e-6, specific_heat=1.0, viscosity=(1e-9, 1e-6),
                 limits=(-1e4, 1e4), step=1, radius=2.1):
    """Switch the heat conduction (CalcTemperatureFactor) on for a case: the parameter-file fields
    it reads (conductivity, resistance, specific heat, viscosity A T + B, temperature limits,
    ThermalStep, radius coefficient), the initial particle temperatures and, for every non-drag
    wall model, its temperature, conductivity and resistivity.  Returns the case."""
    p = case.params
    p.bUpdateTemperature = 1
    p.ThermalStep = step
    p.TemperatureInfluenceRadiusCoefficient = radius
    p.FruidThermalConductivity = conductivity
    p.FruidThermalResistance = resistance
    p.FruidSpecificHeat = specific_heat
    p.ViscosityUpdateParameterA, p.ViscosityUpdateParameterB = viscosity
    p.MinWallTemperature, p.MaxWallTemperature = limits
    if temperature is not None:
        case.particles.temperature[:] = temperature
    for m in case.models:
        if not m.isDrag:
            if wall_temperature is not None:
                m.Temperature = wall_temperature
            m.WallThermalConductivity = wall_conductivity
            m.WallThermalResistivity = wall_resistivity
    return case
