"""Synthetic scenes used by the tests and the benchmark.

All scenes share dx = 0.01, h = 2 dx = 0.02 (cell = h), rest density 1000,
m = rho0 dx^3, Tait gamma = 7 and K = c0^2 with c0 = 10 x the 5 m/s velocity
cap (c0 = 50, K = 2500), XSPH 0.5, dt = 1e-4 with one substep, vdamp 0.5,
limits 5 / 150, particle size dx / 2, gravity (0, -9.8, 0), and the "plain"
coefficients adhesion = viscosity = surface tension = 0 (FluidSystem::setup,
libfluid/cudasrc/fluidsystem.cu:137-140).  Positions come from SplitMix64
seeded generators in the library, so every consumer sees identical bytes.

  dambreak_c1   32^3 jittered lattice (x,y,z in [0.005, 0.315]) in a
                1.0 x 0.5 x 0.5 box, grid 50 x 25 x 25; 32,768 particles
  column_c2     128 x 64 x 128 jittered lattice (1.28 x 0.64 x 1.28 m) in the
                corner of a 2.56 x 1.28 x 2.56 box (4x its footprint), grid
                128 x 64 x 128; exactly 1,048,576 particles
  spray_c3      256 x 115 x 256 jittered block plus 851,968 uniform spray
                particles in y in [1.5, 2.5) over a 2.56^3 box, grid 128^3;
                8,388,608 particles
  double_dam    the reference double dam break (test/src/myfluidsystem.cpp:6-30)
                on the unit cube with res^3 cells, coefficients a/v/s/d of the
                reference driver (test/src/main.cpp:86-89)
  batched_c4    256 independent scenes: particle counts log-uniform in
                [2048, 65536], each a jittered block in the corner of its own
                box (1.5-2.5x the block per axis) under gravity 9.8 along its
                own random direction; dt 1e-4 (NumPy PCG64 from the seed)
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .scene import LfParams, default_params, fluid_block, jittered_lattice, uniform_points

DX = 0.01
H = 2 * DX
RHO0 = 1000.0
C0 = 10 * 5.0


@dataclasses.dataclass
class Config:
    name: str
    params: LfParams
    pos: np.ndarray
    vel: np.ndarray
    dt: float
    steps: int

    @property
    def n(self) -> int:
        return self.pos.shape[0]


def _plain_params(box, grid) -> LfParams:
    p = default_params(1)
    p.rest_density = RHO0
    p.gas_constant = C0 * C0
    p.gamma = 7.0
    p.mass = RHO0 * DX ** 3
    p.particle_size = DX / 2
    p.xsph_scale = 0.5
    p.surface_tension = 0.0
    p.viscosity = 0.0
    p.adhesion = 0.0
    p.vel_limit = 5.0
    p.acc_limit = 150.0
    p.vdamp = 0.5
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.8, 0.0
    p.h = np.float32(H)
    inv = np.float32(1.0) / np.float32(H)
    for a in range(3):
        p.inv_cell[a] = inv
        p.grid[a] = grid[a]
        p.box[a] = box[a]
    return p


def dambreak_c1(seed: int = 1) -> Config:
    p = _plain_params((1.0, 0.5, 0.5), (50, 25, 25))
    pos = jittered_lattice(32, 32, 32, DX, (0.0, 0.0, 0.0), 0.05 * DX, seed)
    return Config("dambreak_c1", p, pos, np.zeros_like(pos), 1e-4, 100)


def column_c2(seed: int = 2) -> Config:
    p = _plain_params((2.56, 1.28, 2.56), (128, 64, 128))
    pos = jittered_lattice(128, 64, 128, DX, (0.0, 0.0, 0.0), 0.05 * DX, seed)
    return Config("column_c2", p, pos, np.zeros_like(pos), 1e-4, 1000)


def spray_c3(seed: int = 3) -> Config:
    p = _plain_params((2.56, 2.56, 2.56), (128, 128, 128))
    block = jittered_lattice(256, 115, 256, DX, (0.0, 0.0, 0.0), 0.05 * DX, seed)
    spray = uniform_points(851_968, (0.0, 1.5, 0.0), (2.56, 2.5, 2.56), seed + 30)
    pos = np.concatenate([block, spray], axis=0)
    return Config("spray_c3", p, pos, np.zeros_like(pos), 1e-4, 500)


def double_dam(n: int = 4096, res: int = 16, adhesion: float = 0.1, viscosity: float = 0.2,
               tension: float = 1.0, vdamp: float = 0.5, dt: float = 0.002) -> Config:
    """MyFluidSystem::setupDoubleDamBreak through createFluidBlock semantics."""
    p = default_params(res)
    p.adhesion, p.viscosity, p.surface_tension, p.vdamp = adhesion, viscosity, tension, vdamp
    a, m1, s1 = fluid_block(n // 2, (0.0, 0.0, 0.0), (0.33, 0.9, 0.33), p.rest_density)
    b, m2, s2 = fluid_block(n - n // 2, (0.66, 0.0, 0.66), (1.0, 0.9, 1.0), p.rest_density)
    pos = np.zeros((n, 3), np.float32)  # unfilled slots stay at the origin, as in init()
    pos[: a.shape[0]] = a
    pos[n // 2: n // 2 + b.shape[0]] = b
    p.mass, p.particle_size = m2, s2  # the last block wins (fluidsystem.cu:105,108)
    return Config("double_dam", p, pos, np.zeros_like(pos), dt, 500)


def scaled_c1(nx: int, ny: int, nz: int, seed: int = 7) -> Config:
    """A smaller jittered block in the C1 box (fast parity cases)."""
    p = _plain_params((1.0, 0.5, 0.5), (50, 25, 25))
    pos = jittered_lattice(nx, ny, nz, DX, (0.0, 0.0, 0.0), 0.05 * DX, seed)
    return Config(f"block_{nx}x{ny}x{nz}", p, pos, np.zeros_like(pos), 1e-4, 10)


def dense_block(nx: int = 30, ny: int = 30, nz: int = 30, spacing: float = 0.25 * H, seed: int = 11) -> Config:
    """A finer lattice (default spacing h/4, about 64 particles a cell) at rest
    density (m = rho0 spacing^3, particle size spacing/2), offset from the
    cell grid so that stencil rows inside hold more than 64 slots and rows at
    its faces fewer; in the C1 box."""
    p = _plain_params((1.0, 0.5, 0.5), (50, 25, 25))
    p.mass = RHO0 * spacing ** 3
    p.particle_size = spacing / 2
    pos = jittered_lattice(nx, ny, nz, spacing, (0.503, 0.107, 0.111), 0.05 * spacing, seed)
    return Config(f"dense_{nx}x{ny}x{nz}", p, pos, np.zeros_like(pos), 1e-4, 10)


@dataclasses.dataclass
class BatchConfig:
    name: str
    scenes: list  # Config per scene

    @property
    def n(self) -> int:
        return sum(s.n for s in self.scenes)

    @property
    def cells(self) -> int:
        return sum(s.params.ncells for s in self.scenes)


def block_scene(n: int, box_factor, gravity_dir, seed: int, name: str = "block") -> Config:
    """n particles: the first n points of a jittered lattice (x outermost),
    nearly cubic, in the corner of a box box_factor[a] times the block's extent
    along each axis, grid ceil(box / h) cells; gravity 9.8 along gravity_dir."""
    side = int(np.ceil(n ** (1.0 / 3.0)))
    nx = nz = side
    ny = int(np.ceil(n / (nx * nz)))
    ext = (nx * DX, ny * DX, nz * DX)
    box = [float(np.float32(ext[a] * box_factor[a])) for a in range(3)]
    grid = [int(np.ceil(box[a] / H)) for a in range(3)]
    p = _plain_params(box, grid)
    g = np.asarray(gravity_dir, np.float64)
    g = 9.8 * g / np.linalg.norm(g)
    p.gravity[0], p.gravity[1], p.gravity[2] = (float(v) for v in g)
    pos = jittered_lattice(nx, ny, nz, DX, (0.0, 0.0, 0.0), 0.05 * DX, seed)[:n].copy()
    return Config(name, p, pos, np.zeros_like(pos), 1e-4, 200)


def batched_c4(seed: int = 4, scenes: int = 256, lo: int = 2048, hi: int = 65536) -> BatchConfig:
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(scenes):
        n = int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))
        box_factor = rng.uniform(1.5, 2.5, size=3)
        gdir = rng.normal(size=3)
        out.append(block_scene(n, box_factor, gdir, seed * 1000 + k, name=f"c4_scene{k}"))
    return BatchConfig(f"batched_c4_{scenes}", out)


def spray_only(n: int = 262_144, seed: int = 7) -> Config:
    """Spray alone: n points uniform in the middle of a 1.28 m box (grid
    64^3), a few neighbours each, so near-isolated particles are everywhere."""
    p = _plain_params((1.28, 1.28, 1.28), (64, 64, 64))
    pos = uniform_points(n, (0.1, 0.1, 0.1), (1.18, 1.18, 1.18), seed)
    return Config("spray_only", p, pos, np.zeros_like(pos), 1e-4, 150)


CONFIGS = {"c1": dambreak_c1, "c2": column_c2, "c3": spray_c3, "double_dam": double_dam}


def cloud(ppc: float = 18.0, side: float = 0.2, seed: int = 5) -> Config:
    """A uniform random cloud of about ppc particles per cell in a cube of
    the given side inside the C1 box: stencil rows of 33 to 64 slots whose
    neighbours fall unevenly into their two 32-slot halves, so some
    particles' flattened-walk tables hold exactly 16 entries."""
    p = _plain_params((1.0, 0.5, 0.5), (50, 25, 25))
    n = int(ppc * (side / H) ** 3)
    pos = uniform_points(n, (0.3, 0.1, 0.1), (0.3 + side, 0.1 + side, 0.1 + side), seed)
    return Config(f"cloud_{n}", p, pos, np.zeros_like(pos), 1e-4, 20)
