"""Parameters and seeded synthetic scenes for the SPH/PCISPH workloads.

Scenes are jittered cubic lattices: index order x fastest, then y, then z;
particle (i, j, k) sits at origin + (i + 1/2, j + 1/2, k + 1/2) s plus a
uniform jitter in [-j s, +j s] per component (3 draws per particle in index
order) from numpy's PCG64 seeded with config_id * 1000 + scene_id.

Workloads (sizes as configured):
  wcsph_dam_break   32^3 = 32,768 particles, s = 0.01 m, block in the corner
                    of a 0.64^3 m box; rho0 = 1000, m = rho0 s^3, h = 2s,
                    mu = 0.01, Tait (k = 3, gamma = 7), g = -9.81 y,
                    dt = 1e-4, box clamp with restitution 0.5; 64^3 cells of s.
  pcisph_dam_break  64^3 = 262,144 particles, s = 0.005 m, block at the
                    x-min end of a 1.28 x 0.64 x 0.64 m box; h = 2s, dt = 5e-4,
                    tolerance 1 %, 3..50 iterations, template delta,
                    256 x 128 x 128 cells of s (hash < 2^22).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import (EOS_LINEAR, EOS_TAIT, MODE_CONFIG, MODE_REFERENCE, SOLVER_PCISPH,
                   SOLVER_WCSPH, WALL_CLAMP, WALL_LATTICE, WARM_NONE, Derived, GridDesc,
                   Params, check, lib)


def default_params() -> Params:
    """Reference defaults (SphParticle.cpp:7-22, SphSystem.cpp:9,59,88,196)."""
    p = Params()
    check(lib().psb_default_params(C.byref(p)))
    return p


def derive(p: Params) -> Derived:
    d = Derived()
    check(lib().psb_derive(C.byref(p), C.byref(d)))
    return d


def grid_desc(lo, hi, cell: float) -> GridDesc:
    g = GridDesc()
    for a in range(3):
        g.min[a] = float(lo[a])
        g.dims[a] = int(math.ceil((hi[a] - lo[a]) / cell - 1e-9))
    g.cell_size = float(cell)
    return g


def lattice(n_side, s: float, origin=(0.0, 0.0, 0.0), jitter: float = 0.01,
            seed: int = 0) -> np.ndarray:
    """(N, 4) float32 positions of an nx*ny*nz jittered lattice (w = 0)."""
    nx, ny, nz = (n_side, n_side, n_side) if np.isscalar(n_side) else n_side
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    pos = (base + 0.5) * s + np.asarray(origin, dtype=np.float64)
    if jitter:
        rng = np.random.Generator(np.random.PCG64(seed))
        pos = pos + rng.uniform(-jitter * s, jitter * s, size=pos.shape)
    out = np.zeros((pos.shape[0], 4), dtype=np.float32)
    out[:, :3] = pos.astype(np.float32)
    return out


def lattice_mass(h: float, s: float, rho0: float) -> float:
    """Mass that gives exactly rho0 at an interior particle of a perfect cubic
    lattice of spacing s under the poly6 kernel of support h."""
    c = 315.0 / (64.0 * math.pi * h ** 9)
    n = int(math.ceil(h / s))
    acc = 0.0
    for a in range(-n, n + 1):
        for b in range(-n, n + 1):
            for e in range(-n, n + 1):
                d2 = (a * a + b * b + e * e) * s * s
                if d2 <= h * h:
                    acc += (h * h - d2) ** 3
    return rho0 / (c * acc)


@dataclass
class Scene:
    name: str
    params: Params
    pos4: np.ndarray
    vel4: np.ndarray = None
    vh4: np.ndarray = None
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        n = self.pos4.shape[0]
        if self.vel4 is None:
            self.vel4 = np.zeros((n, 4), np.float32)
        if self.vh4 is None:
            self.vh4 = np.zeros((n, 4), np.float32)

    @property
    def n(self) -> int:
        return int(self.pos4.shape[0])


def wcsph_dam_break(n_side: int = 32, s: float = 0.01, mode: int = MODE_CONFIG,
                    scene_id: int = 0, cell: float | None = None) -> Scene:
    """Plain WCSPH dam break (32^3 by default)."""
    p = default_params()
    p.mode = mode
    p.solver = SOLVER_WCSPH
    p.eos = EOS_TAIT if mode == MODE_CONFIG else EOS_LINEAR
    p.rest_density = 1000.0
    p.mass = 1000.0 * s ** 3
    p.support_radius = 2 * s
    p.gas_stiffness = 3.0
    p.tait_gamma = 7.0
    p.viscosity = 0.01
    p.surface_tension = 0.0 if mode == MODE_CONFIG else 0.0728
    p.particle_radius = 0.5 * s
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.81, 0.0
    p.dt = 1e-4
    L = 2 * n_side * s
    p.box_enabled = 1
    for a in range(3):
        p.box_min[a] = 0.0
        p.box_max[a] = L
    p.elast = 0.5
    p.grid = grid_desc((0.0, 0.0, 0.0), (L, L, L), cell or s)
    p.max_neighbors = 96
    p.lattice_spacing = s
    pos = lattice(n_side, s, seed=1000 + scene_id)
    return Scene("wcsph_dam_break", p, pos, meta={"s": s, "n_side": n_side})


def pcisph_dam_break(n_side=64, s: float = 0.005, mode: int = MODE_CONFIG,
                     scene_id: int = 0, box=(1.28, 0.64, 0.64), tolerance: float = 0.01,
                     min_it: int = 3, max_it: int = 50, warm_start: int = WARM_NONE,
                     calibrate_mass: bool = True, lattice_walls: bool = False,
                     relaxation: float = 1.0, retry_levels: int | None = None) -> Scene:
    """Plain PCISPH dam break (64^3 block at the x-min end of the box).

    lattice_walls: the box walls also act as implicit boundary particles on
    the fluid lattice continued past them (config mode); otherwise they act
    only through the box collider.
    retry_levels: a step whose corrector ends above the tolerance is redone
    as two steps of dt/2, recursively this many levels deep (default 2 in
    config mode, 0 in reference mode).  At dt = 5e-4 the corrector diverges
    at the wall impacts of the 262,144-particle dam break (errors up to
    115 % after 50 iterations); the same states converge at dt/2.
    """
    p = default_params()
    p.mode = mode
    p.solver = SOLVER_PCISPH
    p.eos = EOS_LINEAR
    p.warm_start = warm_start
    h = 2 * s
    p.rest_density = 1000.0
    p.support_radius = h
    p.mass = lattice_mass(h, s, 1000.0) if calibrate_mass else 1000.0 * s ** 3
    p.gas_stiffness = 3.0
    p.viscosity = 0.01
    p.surface_tension = 0.0 if mode == MODE_CONFIG else 0.0728
    p.particle_radius = 0.5 * s
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.81, 0.0
    p.dt = 5e-4
    p.box_enabled = 1
    for a in range(3):
        p.box_min[a] = 0.0
        p.box_max[a] = float(box[a])
    p.elast = 0.5
    p.grid = grid_desc((0.0, 0.0, 0.0), box, s)
    p.max_neighbors = 96
    p.tolerance = tolerance
    p.min_iterations = min_it
    p.max_iterations = max_it
    p.lattice_spacing = s
    p.delta_local_cap = 1
    p.wall_model = WALL_LATTICE if lattice_walls else WALL_CLAMP
    p.relaxation = relaxation
    p.retry_levels = (2 if mode == MODE_CONFIG else 0) if retry_levels is None else retry_levels
    pos = lattice(n_side, s, seed=2000 + scene_id)
    return Scene("pcisph_dam_break", p, pos, meta={"s": s, "n_side": n_side})


def batched_scenes(num_scenes: int = 64, s: float = 0.005, seed: int = 5000,
                   min_count: int = 4096, max_count: int = 131072,
                   warm_start: int = WARM_NONE, tolerance: float = 0.01, min_it: int = 3,
                   max_it: int = 50, lattice_walls: bool = False) -> list:
    """Independent PCISPH scenes for batched serving.

    Particle counts are drawn log-uniform in [min_count, max_count]
    (numpy PCG64, `seed`).  Even scenes are dam breaks: a column a x h x a
    (a = round((N/2)^(1/3)), h = ceil(N / a^2) layers, the last one partly
    filled) at the x-min end of a 4a x max(2.5a, 1.25h) x a box.  Odd scenes
    are drops: a cube (a = round(N^(1/3))) released at mid height of a
    3a x 3a x 3a box.  Counts are exact: the lattice is filled in x-fastest
    order and cut at N.  Every scene shares the material of the PCISPH dam
    break (spacing s, h = 2s, calibrated mass, 1% tolerance); its params
    differ only in box and grid.  Jitter seed of scene i: seed + i.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = np.floor(np.exp(rng.uniform(np.log(min_count), np.log(max_count), num_scenes)))
    out = []
    for i, N in enumerate(counts.astype(np.int64)):
        N = int(N)
        if i % 2 == 0:
            a = max(2, int(round((N / 2) ** (1.0 / 3.0))))
            hgt = -(-N // (a * a))
            box = (4 * a * s, max(2.5 * a, 1.25 * hgt) * s, a * s)
            origin = (0.0, 0.0, 0.0)
            dims = (a, hgt, a)
            kind = "dam_break"
        else:
            a = max(2, int(round(N ** (1.0 / 3.0))))
            hgt = -(-N // (a * a))
            box = (3 * a * s, max(3 * a, 2 * hgt) * s, 3 * a * s)
            origin = (a * s, 0.5 * (box[1] - hgt * s), a * s)
            dims = (a, hgt, a)
            kind = "drop"
        pos = lattice(dims, s, origin=origin, seed=seed + i)[:N]
        sc = pcisph_dam_break(n_side=1, s=s, box=box, tolerance=tolerance, min_it=min_it,
                              max_it=max_it, warm_start=warm_start, lattice_walls=lattice_walls)
        sc.params.grid = grid_desc((0.0, 0.0, 0.0), box, s)
        out.append(Scene(f"{kind}_{i}", sc.params, pos, meta={"kind": kind, "box": box, "s": s}))
    return out


def double_dam_break_dem(block=(128, 128, 64), s: float = 0.005, box=(2.56, 1.28, 1.28),
                         num_spheres: int = 512, sphere_radius_cells: float = 3.0,
                         sphere_density: float = 500.0, spring: float = 1e4,
                         damping_ratio: float = 0.3, seed: int = 3000, tolerance: float = 0.01,
                         min_it: int = 3, max_it: int = 50, warm_start: int = WARM_NONE,
                         lattice_walls: bool = True, relaxation: float = 1.0,
                         dt: float = 2.5e-4) -> Scene:
    """PCISPH double dam break with rigid DEM spheres.

    Two fluid blocks of block = (nx, ny, nz) lattice particles at the x-min
    and x-max ends of the box (2 x 128 x 128 x 64 = 2,097,152 particles at
    full size), walls as lattice boundary particles, and `num_spheres`
    spheres of radius sphere_radius_cells * s and the given density placed
    uniformly at random (PCG64 `seed`) in the free space between the blocks,
    at least 2 radii + s apart and one radius + s from the walls.  Contact
    spring k_n = spring; the normal dashpot coefficient is
    c = 2 damping_ratio sqrt(k_n m / 2) (m / 2: reduced mass of two spheres).
    The corrector runs the full template delta (relaxation 1) with failed
    steps redone as substeps (retry_levels 2).  dt = 2.5e-4: the two 0.64 m
    columns meet head-on at ~5 m/s each around t = 0.22 s, with the spheres
    carried at up to 3.6 m/s; measured on MI355X, at 2.5e-4 every one of 1200
    steps (t = 0.3 s, through the collision) ends within the 1 % tolerance
    at 8-10 iterations per step and no retried step, while at 5e-4 the scene
    with spheres fails from the collision on (worst 23 % over 1000 steps,
    31 ms per step) -- without spheres it holds at 5e-4 too, by retries
    (229 of 1000 steps redone).  meta holds the sphere centres and
    parameters.
    """
    nx, ny, nz = block
    sc = pcisph_dam_break(n_side=1, s=s, box=box, tolerance=tolerance, min_it=min_it,
                          max_it=max_it, warm_start=warm_start, lattice_walls=lattice_walls,
                          relaxation=relaxation)
    p = sc.params
    p.dt = dt
    p.grid = grid_desc((0.0, 0.0, 0.0), box, s)
    left = lattice((nx, ny, nz), s, seed=seed)
    right = lattice((nx, ny, nz), s, origin=(box[0] - nx * s, 0.0, 0.0), seed=seed + 1)
    pos = np.concatenate([left, right])
    R = sphere_radius_cells * s
    mass = sphere_density * 4.0 / 3.0 * math.pi * R ** 3
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    lo = np.array([nx * s + R + s, R + s, R + s])
    hi = np.array([box[0] - nx * s - R - s, box[1] - R - s, box[2] - R - s])
    centers = []
    tries = 0
    while len(centers) < num_spheres:
        tries += 1
        if tries > 1000 * max(1, num_spheres):
            raise ValueError("spheres do not fit")
        c = rng.uniform(lo, hi)
        if all(np.sum((c - q) ** 2) >= (2 * R + s) ** 2 for q in centers):
            centers.append(c)
    centers = np.array(centers, np.float32).reshape(-1, 3)
    meta = {"s": s, "block": block, "sphere_centers": centers, "sphere_radius": R,
            "sphere_density": sphere_density, "spring": spring,
            "damping": 2.0 * damping_ratio * math.sqrt(spring * mass / 2.0),
            "sphere_mass": mass}
    return Scene("double_dam_break_dem", p, pos, meta=meta)


# The reference's fluid materials (Ressources/FluidMaterial/Water.xml and
# Mucus.xml): particle_mass, supportRadius, restDensity, gasStiffness,
# viscosity, surfaceTension, threshold.
WATER = dict(mass=0.02, support_radius=0.0457, rest_density=998.29, gas_stiffness=3.0,
             viscosity=3.5, surface_tension=0.0728, threshold=7.065)
MUCUS = dict(mass=0.04, support_radius=0.0726, rest_density=1000.0, gas_stiffness=5.0,
             viscosity=36.0, surface_tension=6.0, threshold=5.0)


def material(base: Params, **fields) -> Params:
    """A copy of `base` with the given material fields."""
    p = base.copy()
    for k, v in fields.items():
        setattr(p, k, v)
    return p


def two_fluid_scene(solver: int = SOLVER_PCISPH, n_side: int = 8, s: float = 0.03,
                    seed: int = 0, dt: float = 0.005, rest_spacing: bool = False) -> Scene:
    """Two touching blocks in reference mode, water (x-min block) and mucus,
    each n_side^3 on a lattice of spacing s, in a box of at least 1.0 x 0.6 x
    0.6 m (grown to hold larger blocks with 0.1 m to spare).  rest_spacing:
    each block on its material's own spacing (m / rho0)^(1/3) instead, so
    that both start near their rest density.

    The system's parameters carry the larger support radius (mucus, which
    sizes the neighbour search); meta["materials"] holds [water, mucus] and
    meta["index"] each particle's material (0 for the first block)."""
    p = default_params()
    p.mode = MODE_REFERENCE
    p.solver = solver
    p.eos = EOS_LINEAR
    for k, v in MUCUS.items():
        setattr(p, k, v)
    p.particle_radius = 0.5 * s
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.81, 0.0
    p.dt = dt
    o = 0.05
    sa, sb = s, s
    if rest_spacing:
        sa = (WATER["mass"] / WATER["rest_density"]) ** (1.0 / 3.0)
        sb = (MUCUS["mass"] / MUCUS["rest_density"]) ** (1.0 / 3.0)
    smax = max(sa, sb)
    box = (max(1.0, o + n_side * (sa + sb) + 0.1), max(0.6, o + n_side * smax + 0.1),
           max(0.6, o + n_side * smax + 0.1))
    p.box_enabled = 1
    for a in range(3):
        p.box_min[a] = 0.0
        p.box_max[a] = box[a]
    p.elast = 0.0
    p.grid = grid_desc((0.0, 0.0, 0.0), box, MUCUS["support_radius"] / 2)
    p.max_neighbors = 160
    p.lattice_spacing = s
    p.min_iterations = p.max_iterations = 3
    a = lattice(n_side, sa, origin=(o, o, o), jitter=0.02, seed=2000 + seed)
    b = lattice(n_side, sb, origin=(o + n_side * sa, o, o), jitter=0.02, seed=3000 + seed)
    pos = np.concatenate([a, b])
    index = np.repeat(np.arange(2, dtype=np.uint8), [a.shape[0], b.shape[0]])
    mats = [material(p, **WATER), material(p, **MUCUS)]
    return Scene("two_fluid", p, pos, meta={"materials": mats, "index": index, "s": s})
