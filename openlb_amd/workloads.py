"""Seeded synthetic workloads for the block-lattice time step.

Every generator returns a :class:`Case`: the per-cell dynamics ids, the
initial moments and the collision parameters. The layout is core-only,
shape ``(nx, ny, nz)``, with z contiguous. This matches OpenLB's x-major cell order
(``src/core/blockStructure.h:65-78``).

Dynamics ids follow the backend's convention. 0 is NoDynamics, 1 is the lattice's bulk
collision, 2 is BounceBack (``src/dynamics/dynamics.h:472-483``), and 3..255 are
BounceBackVelocity walls (``dynamics.h:550-555``), each with its own (rho, u).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BULK_KINDS = ("bgk", "forced_bgk", "trt", "forced_trt", "smagorinsky")


@dataclasses.dataclass
class Case:
    name: str
    q: int
    kind: str
    shape: tuple[int, int, int]
    periodic: tuple[int, int, int]
    ids: np.ndarray            # uint8 (nx, ny, nz)
    rho_u: np.ndarray          # float64 (4, nx, ny, nz)
    omega: float
    magic: float = 0.25
    smago: float = 0.0
    force: tuple[float, float, float] = (0.0, 0.0, 0.0)
    walls: dict = dataclasses.field(default_factory=dict)   # id -> (rho, (ux, uy, uz))
    local: dict = dataclasses.field(default_factory=dict)   # id -> "velocity" | "pressure"
    force_field: np.ndarray | None = None  # float64 (3, nx, ny, nz): per-cell FORCE of the forced kinds

    @property
    def ncore(self) -> int:
        return int(np.prod(self.shape))

    @property
    def n_active(self) -> int:
        """Cells with dynamics (OpenLB's getNvoxel(): material != 0)."""
        return int(np.count_nonzero(self.ids))


def _grid(shape):
    return np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")


def taylor_green(shape=(64, 64, 64), u0=0.05, tau=0.6, q=19):
    """3D Taylor-Green vortex on a fully periodic lattice.

    The velocity is u = u0*(sin x cos y cos z, -cos x sin y cos z, 0), with
    x = 2*pi*i/N, and rho = 1. This is the initial field of
    ``examples/turbulence/tgv3d/tgv3d.cpp:126-137`` on lattice coordinates.
    """
    x, y, z = _grid(shape)
    kx, ky, kz = (2 * math.pi / n for n in shape)
    rho_u = np.empty((4,) + tuple(shape))
    rho_u[0] = 1.0
    rho_u[1] = u0 * np.sin(kx * x) * np.cos(ky * y) * np.cos(kz * z)
    rho_u[2] = -u0 * np.cos(kx * x) * np.sin(ky * y) * np.cos(kz * z)
    rho_u[3] = 0.0
    ids = np.ones(shape, dtype=np.uint8)
    return Case("taylor_green", q, "bgk", tuple(shape), (1, 1, 1), ids, rho_u, 1.0 / tau)


def place_spheres(shape, seed, solid_fraction=0.1, rmin=3, rmax=6, region=None):
    """Seeded solid spheres until the solid share of ``region`` reaches ``solid_fraction``.

    Centres are uniform over the region and radii are uniform integers in [rmin, rmax].
    A cell is solid when its centre lies within the radius. Spheres wrap
    periodically in x and are clipped in y and z. This is the build's reading of
    "seeded Bernoulli(0.1) spheres": placement stops at the first sphere that
    takes the solid fraction to at least 0.1.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz = shape
    solid = np.zeros(shape, dtype=bool)
    if region is None:
        region = np.ones(shape, dtype=bool)
    target = solid_fraction * np.count_nonzero(region)
    count = 0
    while count < target:
        r = int(rng.integers(rmin, rmax + 1))
        cx = rng.uniform(0, nx)
        cy = rng.uniform(0, ny)
        cz = rng.uniform(0, nz)
        xs = np.arange(int(math.floor(cx - r)), int(math.ceil(cx + r)) + 1)
        ys = np.arange(max(0, int(math.floor(cy - r))), min(ny, int(math.ceil(cy + r)) + 1))
        zs = np.arange(max(0, int(math.floor(cz - r))), min(nz, int(math.ceil(cz + r)) + 1))
        dx = (xs - cx)[:, None, None]
        dy = (ys - cy)[None, :, None]
        dz = (zs - cz)[None, None, :]
        inside = dx * dx + dy * dy + dz * dz <= r * r
        xi = np.mod(xs, nx)
        sub = solid[np.ix_(xi, ys, zs)]
        new = inside & ~sub & region[np.ix_(xi, ys, zs)]
        count += int(np.count_nonzero(new))
        solid[np.ix_(xi, ys, zs)] = sub | (inside & region[np.ix_(xi, ys, zs)])
    return solid


def porous_channel(shape=(256, 256, 256), seed=20261202, tau=0.55, force=1e-6,
                   solid_fraction=0.1, rmin=3, rmax=6, q=19, kind="forced_bgk"):
    """Porous channel: periodic in x, one-cell BounceBack walls at the y/z bounds,
    seeded BounceBack spheres in the interior, Guo body force along x, rho=1, u=0."""
    nx, ny, nz = shape
    ids = np.ones(shape, dtype=np.uint8)
    wall = np.zeros(shape, dtype=bool)
    wall[:, 0, :] = wall[:, -1, :] = True
    wall[:, :, 0] = wall[:, :, -1] = True
    spheres = place_spheres(shape, seed, solid_fraction, rmin, rmax, region=~wall)
    ids[wall | spheres] = 2
    rho_u = np.zeros((4,) + tuple(shape))
    rho_u[0] = 1.0
    return Case("porous_channel", q, kind, tuple(shape), (1, 0, 0), ids, rho_u,
                1.0 / tau, force=(force, 0.0, 0.0))


def lid_driven_cavity(shape=(32, 32, 32), tau=0.6, lid_u=0.05, q=19, kind="bgk"):
    """Lid-driven cavity. Every boundary cell is BounceBack, except the top (y = ny-1)
    face, which is a BounceBackVelocity wall (id 3) moving along x at rho = 1.
    This is the bounce-back variant of ``examples/laminar/cavity3dBenchmark/cavity3d.cpp:59-87``."""
    ids = np.ones(shape, dtype=np.uint8)
    ids[0, :, :] = ids[-1, :, :] = 2
    ids[:, 0, :] = 2
    ids[:, :, 0] = ids[:, :, -1] = 2
    ids[1:-1, -1, 1:-1] = 3
    rho_u = np.zeros((4,) + tuple(shape))
    rho_u[0] = 1.0
    return Case("lid_driven_cavity", q, kind, tuple(shape), (0, 0, 0), ids, rho_u,
                1.0 / tau, walls={3: (1.0, (lid_u, 0.0, 0.0))})


def couette_channel(shape=(16, 12, 10), tau=0.7, wall_u=0.05, force=1e-5, q=19, kind="forced_bgk"):
    """Plane Couette-Poiseuille channel, periodic in x and z: BounceBack at y = 0,
    a BounceBackVelocity wall (id 3) moving along x at y = ny-1, and a uniform
    body force along x on the fluid."""
    ids = np.ones(shape, dtype=np.uint8)
    ids[:, 0, :] = 2
    ids[:, -1, :] = 3
    rho_u = np.zeros((4,) + tuple(shape))
    rho_u[0] = 1.0
    forced = kind in ("forced_bgk", "forced_trt")
    return Case("couette_channel", q, kind, tuple(shape), (1, 0, 1), ids, rho_u, 1.0 / tau,
                force=(force if forced else 0.0, 0.0, 0.0), walls={3: (1.0, (wall_u, 0.0, 0.0))})


def cavity_benchmark(size=100, q=19):
    """The cavity of OpenLB's benchmark harness in its bounce-back-lid variant
    (examples/laminar/cavity3dBenchmark/cavity3d.cpp:39-57, 85-87, 154-250):
    (size+1)^3 cells, BounceBack boundary, the whole top face (y = size) a
    BounceBackVelocity lid at lattice velocity 0.01, BGK with
    tau = 1/2 + 3 * 1e-5 * size (lattice velocity 0.01, Re 1000)."""
    n = size + 1
    ids = np.full((n, n, n), 2, dtype=np.uint8)
    ids[1:-1, 1:-1, 1:-1] = 1
    ids[:, -1, :] = 3
    rho_u = np.zeros((4, n, n, n))
    rho_u[0] = 1.0
    tau = 0.5 + 3.0 * 1e-5 * size
    return Case("cavity_benchmark", q, "bgk", (n, n, n), (0, 0, 0), ids, rho_u, 1.0 / tau,
                walls={3: (1.0, (0.01, 0.0, 0.0))})


def local_velocity_cavity(shape=(32, 32, 32), tau=0.6, lid_u=0.05, q=19, boundary="velocity"):
    """Lid-driven cavity with OpenLB's default boundary of the benchmark harness:
    LocalVelocity (regularized) boundaries on the whole box, at rest on the walls
    (material 2) and moving along x on the lid (material 3).  The lid is the top
    face (y = ny-1) without its edges and corners, the cells OpenLB's
    ``rename(2, 3, 1, lid)`` selects (cavity3d.cpp:39-57, 68-76, 113-131;
    blockGeometry.hh:537-565).  boundary="interp_velocity" uses OpenLB's
    InterpolatedVelocity instead (src/boundary/interpolatedVelocity3D.h)."""
    ids = np.full(shape, 2, dtype=np.uint8)
    ids[1:-1, 1:-1, 1:-1] = 1
    ids[1:-1, -1, 1:-1] = 3
    rho_u = np.zeros((4,) + tuple(shape))
    rho_u[0] = 1.0
    name = "local_velocity_cavity" if boundary == "velocity" else "interpolated_velocity_cavity"
    return Case(name, q, "bgk", tuple(shape), (0, 0, 0), ids, rho_u, 1.0 / tau,
                walls={2: (1.0, (0.0, 0.0, 0.0)), 3: (1.0, (lid_u, 0.0, 0.0))},
                local={2: boundary, 3: boundary})


def cavity_benchmark_local(size=100, q=19):
    """OpenLB's benchmark cavity in its default configuration (LID_DRIVEN without
    LID_DRIVEN_BOUNCE_BACK, cavity3d.cpp:32-36, 68-76): (size+1)^3 cells,
    LocalVelocity boundaries, lid velocity 0.01, tau = 1/2 + 3e-5 * size."""
    c = local_velocity_cavity((size + 1,) * 3, tau=0.5 + 3.0 * 1e-5 * size, lid_u=0.01, q=q)
    c.name = "cavity_benchmark_local"
    return c


def velocity_channel(shape=(24, 12, 10), u_in=0.03, rho_out=1.0, tau=0.6, q=19, inlet="velocity",
                     outlet="pressure", periodic_z=False):
    """Duct with a LocalVelocity inlet (x = 0, material 3, uniform u_in along x),
    a LocalPressure outlet (x = nx-1, material 4, density rho_out) and BounceBack
    walls (material 2) on the y and z faces, including the inlet's and outlet's
    rims.  Starts at rest.  inlet="interp_velocity" / outlet="interp_pressure"
    select OpenLB's InterpolatedVelocity / InterpolatedPressure instead.  With
    periodic_z the z faces are periodic instead of walls (a plane channel)."""
    ids = np.full(shape, 2, dtype=np.uint8)
    zs = slice(None) if periodic_z else slice(1, -1)
    ids[:, 1:-1, zs] = 1
    ids[0, 1:-1, zs] = 3
    ids[-1, 1:-1, zs] = 4
    rho_u = np.zeros((4,) + tuple(shape))
    rho_u[0] = 1.0
    return Case("velocity_channel", q, "bgk", tuple(shape), (0, 0, int(periodic_z)), ids, rho_u, 1.0 / tau,
                walls={3: (1.0, (u_in, 0.0, 0.0)), 4: (rho_out, (0.0, 0.0, 0.0))},
                local={3: inlet, 4: outlet})


def step_cavity(shape=(14, 12, 16), block=(6, 5, 7), tau=0.6, lid_u=0.05, q=19):
    """Lid-driven cavity whose lower corner holds a solid block of wall material
    (x < block[0], y < block[1], z < block[2]) with LocalVelocity boundaries on
    every wall: the block's convex edges and corner give OpenLB's internal edge
    and corner boundary types next to the box's flat faces, external edges and
    corners."""
    c = local_velocity_cavity(shape, tau=tau, lid_u=lid_u, q=q)
    bx, by, bz = block
    c.ids[:bx, :by, :bz] = 2
    # the block interior is surrounded by wall cells only: no dynamics there,
    # like a cell OpenLB's innerClean() resets to material 0
    inner = np.zeros(shape, dtype=bool)
    inner[1:bx - 1, 1:by - 1, 1:bz - 1] = True
    c.ids[inner] = 0
    c.name = "step_cavity"
    return c


def cavity_ensemble(count=512, seed=20261204, lo=16, hi=96, tau=(0.52, 1.2), lid=(0.02, 0.1), q=19):
    """Config-4 ensemble: independent lid-driven cavities. Each edge is drawn
    log-uniform in [lo, hi] per axis, tau uniform in ``tau``, and the lid velocity
    uniform in ``lid`` along x (lid_driven_cavity layout: BounceBack box, moving top face)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cases = []
    for k in range(count):
        shape = tuple(int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(3))
        t = float(rng.uniform(*tau))
        u = float(rng.uniform(*lid))
        c = lid_driven_cavity(shape, tau=t, lid_u=u, q=q)
        c.name = f"cavity_{k}"
        cases.append(c)
    return cases


def turbulent_channel(shape=(384, 192, 192), tau=0.505, smago=0.14, u_max=0.05,
                      noise=0.01, seed=20261203, q=27):
    """Smagorinsky channel. It is periodic in x and z, with BounceBack walls at y = 0 and
    y = ny-1. The laminar Poiseuille profile u_x(y) carries seeded uniform
    +/-``noise`` relative velocity noise per component."""
    nx, ny, nz = shape
    ids = np.ones(shape, dtype=np.uint8)
    ids[:, 0, :] = ids[:, -1, :] = 2
    rng = np.random.Generator(np.random.PCG64(seed))
    h = ny - 2
    yc = np.arange(ny, dtype=np.float64) - 0.5
    prof = np.clip(4.0 * u_max * yc * (h - yc) / (h * h), 0.0, None)
    rho_u = np.zeros((4,) + tuple(shape))
    rho_u[0] = 1.0
    rho_u[1] = prof[None, :, None]
    for d in range(1, 4):
        rho_u[d] += u_max * noise * rng.uniform(-1.0, 1.0, size=shape)
    rho_u[1:, ids == 2] = 0.0
    return Case("turbulent_channel", q, "smagorinsky", tuple(shape), (1, 0, 1), ids, rho_u,
                1.0 / tau, smago=smago)


def kolmogorov_flow(shape=(16, 14, 12), amplitude=2e-5, tau=0.62, q=19, kind="forced_bgk", walls=False,
                    seed=3):
    """Forced flow with a FORCE field that differs from cell to cell: a
    Kolmogorov shear force along x varying with y, plus smaller transverse
    components varying with z and x, F(x, y, z) = A (sin(2 pi (y+1/2)/ny),
    cos(2 pi (z+1/2)/nz) / 2, sin(2 pi (x+1/2)/nx) / 4).  Fully periodic, or
    (walls=True) periodic in x and z between BounceBack walls at the y bounds
    with a few seeded BounceBack obstacles.  Starts at rest."""
    x, y, z = _grid(shape)
    nx, ny, nz = shape
    F = np.empty((3,) + tuple(shape))
    F[0] = amplitude * np.sin(2 * math.pi * (y + 0.5) / ny)
    F[1] = 0.5 * amplitude * np.cos(2 * math.pi * (z + 0.5) / nz)
    F[2] = 0.25 * amplitude * np.sin(2 * math.pi * (x + 0.5) / nx)
    ids = np.ones(shape, dtype=np.uint8)
    periodic = (1, 1, 1)
    if walls:
        periodic = (1, 0, 1)
        ids[:, 0, :] = 2
        ids[:, -1, :] = 2
        rng = np.random.Generator(np.random.PCG64(seed))
        for _ in range(3):
            cx, cy, cz = rng.integers(0, nx), rng.integers(2, ny - 2), rng.integers(0, nz)
            ids[cx, cy, cz] = 2
    rho_u = np.zeros((4,) + tuple(shape))
    rho_u[0] = 1.0
    name = "kolmogorov_channel" if walls else "kolmogorov_flow"
    return Case(name, q, kind, tuple(shape), periodic, ids, rho_u, 1.0 / tau, force_field=F)


def random_periodic(shape, seed, q=19, kind="bgk", tau=0.7, u_amp=0.04, rho_amp=0.01,
                    magic=0.25, smago=0.0, force=(0.0, 0.0, 0.0)):
    """Fully periodic box with smooth random moments. This is a parity case whose
    populations carry every direction non-trivially."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x, y, z = _grid(shape)
    rho_u = np.empty((4,) + tuple(shape))
    for comp in range(4):
        acc = np.zeros(shape)
        for _ in range(3):
            kx, ky, kz = (rng.integers(1, 3) * 2 * math.pi / n for n in shape)
            ph = rng.uniform(0, 2 * math.pi, size=3)
            acc += np.sin(kx * x + ph[0]) * np.cos(ky * y + ph[1]) * np.sin(kz * z + ph[2])
        acc /= 3.0
        rho_u[comp] = (1.0 + rho_amp * acc) if comp == 0 else u_amp * acc
    ids = np.ones(shape, dtype=np.uint8)
    return Case("random_periodic", q, kind, tuple(shape), (1, 1, 1), ids, rho_u,
                1.0 / tau, magic=magic, smago=smago, force=tuple(force))
