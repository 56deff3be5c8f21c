"""Seeded dam-break scenes for parity tests and benchmarks.

A scene is a jittered lattice block (the lattice generator of
samples/main.cpp:36-56, x outermost, z innermost, ids in that order) plus the
solver parameters. The jitter is a counter-based hash of (seed, particle id,
axis), so any consumer regenerates bit-identical inputs.

Scaling: every scene is the same dimensionless recipe in solver units (the
reference's own Config::scale, world = solver * scale, oclsph.hpp:301,374,414):
solver spacing 0.05, h = 0.1, mass 1, gravity -9.8, dt 1/120, the regime the
reference's constants (rest density 6378, CFM epsilon 600, a mass-free
lambda gradient, oclsph_kernel.h:20-31,323-331) are tuned for. A world
spacing s is reached with scale = s / 0.05, so h_world = 2 s (~33 lattice
neighbours). Shrinking h in solver units instead would leave epsilon and the
mass-free gradient at the wrong scale and the step diverges.
"""
import dataclasses
from typing import Sequence, Tuple

import numpy as np

REST_DENSITY = 6378.0
SOLVER_SPACING = 0.05
SOLVER_H = 0.1

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)


def hash_uniform(seed, ids, axis):
    """splitmix64 of (seed, id, axis) -> float64 uniform in [0, 1)."""
    with np.errstate(over="ignore"):
        ids = np.asarray(ids, dtype=np.uint64)
        x = (np.uint64(seed) * _M1 + ids * _M2 + np.uint64(axis + 1) * _M3)
        x = (x ^ (x >> np.uint64(30))) * _M2
        x = (x ^ (x >> np.uint64(27))) * _M3
        x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def lattice_block(dims, spacing, origin, jitter, seed, id_offset=0):
    """float32 positions of a dims[0] x dims[1] x dims[2] lattice, x outermost."""
    nx, ny, nz = dims
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    ids = np.arange(idx.shape[0], dtype=np.uint64) + np.uint64(id_offset)
    pos = np.asarray(origin, np.float64) + idx * spacing
    if jitter:
        for k in range(3):
            pos[:, k] += (2.0 * hash_uniform(seed, ids, k) - 1.0) * jitter
    return pos.astype(np.float32), ids


@dataclasses.dataclass
class Scene:
    name: str
    h: float            # solver units
    spacing: float      # world units
    scale: float        # world = solver * scale
    position: np.ndarray
    velocity: np.ndarray
    mass: np.ndarray
    type: np.ndarray
    colour: np.ndarray
    ids: np.ndarray
    min_bound: Tuple[float, float, float]
    max_bound: Tuple[float, float, float]
    iterations: int = 4
    dt: float = 1.0 / 120.0
    gravity: Tuple[float, float, float] = (0.0, -9.8, 0.0)
    xsph_c: float = 0.01
    steps: int = 10
    mc_samples: Sequence[int] = (64, 64, 64)
    isolevel: float = 0.5 * REST_DENSITY
    seed: int = 1

    @property
    def n(self):
        return self.position.shape[0]

    def params(self, flags=0, xsph_c=None, iterations=None, density_tol=0.0):
        from . import pbf_params
        return pbf_params(self.h, dt=self.dt, scale=self.scale,
                          iterations=self.iterations if iterations is None else iterations,
                          gravity=self.gravity, min_bound=self.min_bound,
                          max_bound=self.max_bound,
                          xsph_c=self.xsph_c if xsph_c is None else xsph_c, flags=flags,
                          density_tol=density_tol)

    def mc(self, attributes=0):
        from . import mc_density_params
        return mc_density_params(self.isolevel, tuple(self.mc_samples), attributes)


def _scene(name, blocks, spacing, box, seed, steps, mc_samples, colour=0xFF2D1010):
    pos, ids = [], []
    off = 0
    for dims, origin in blocks:
        p, i = lattice_block(dims, spacing, origin, 0.1 * spacing, seed, off)
        pos.append(p)
        ids.append(i)
        off += p.shape[0]
    pos = np.concatenate(pos)
    n = pos.shape[0]
    return Scene(
        name=name, h=SOLVER_H, spacing=spacing, scale=spacing / SOLVER_SPACING, position=pos,
        velocity=np.zeros((n, 3), np.float32),
        mass=np.ones(n, np.float32),
        type=np.zeros(n, np.int32), colour=np.full(n, colour, np.uint32),
        ids=np.concatenate(ids), min_bound=(0.0, 0.0, 0.0), max_bound=tuple(box),
        steps=steps, mc_samples=mc_samples, seed=seed)


def dam_break_8k(seed=1):
    """8,000 particles: 20^3 lattice, spacing 0.05 from (0.025,..), jitter
    +-0.005, 2x1x1 box, h = 0.1, 10 steps, marching cubes at 64^3."""
    s = 0.05
    return _scene("dam_break_8k", [((20, 20, 20), (s / 2,) * 3)], s, (2.0, 1.0, 1.0), seed, 10,
                  (64, 64, 64))


def dam_break_1m(seed=2):
    """1,048,576 particles: a 128 (x) x 128 (y) x 64 (z) lattice column, world
    spacing 0.0125 (scale 0.25), 4x2x2 box, h = 0.025 world, 100 steps, marching
    cubes at 256^3 every step. (A 128^3 lattice would hold 2,097,152 particles;
    the 1,048,576 count is kept.)"""
    s = 0.0125
    return _scene("dam_break_1m", [((128, 128, 64), (s / 2,) * 3)], s, (4.0, 2.0, 2.0), seed,
                  100, (256, 256, 256))


def double_dam_break_4m(seed=4):
    """4,194,304 particles: two 128^3 blocks, world spacing 0.00625 (scale
    0.125), h = 0.0125 world, at opposite x ends of a 4x2x2 box, 20 steps."""
    s = 0.00625
    side = 128 * s
    return _scene("double_dam_break_4m",
                  [((128, 128, 128), (s / 2,) * 3),
                   ((128, 128, 128), (4.0 - side + s / 2, s / 2, s / 2))],
                  s, (4.0, 2.0, 2.0), seed, 20, (256, 256, 256))


def small_block(n_side=10, spacing=0.05, seed=7, box=(1.0, 1.0, 1.0), jitter=True):
    """A small jittered block for quick parity checks."""
    s = spacing
    p, ids = lattice_block((n_side,) * 3, s, (s / 2 + 0.1,) * 3, 0.1 * s if jitter else 0.0,
                           seed)
    n = p.shape[0]
    return Scene(name="small_block_%d" % n, h=SOLVER_H, spacing=s, scale=s / SOLVER_SPACING,
                 position=p, velocity=np.zeros((n, 3), np.float32),
                 mass=np.ones(n, np.float32),
                 type=np.zeros(n, np.int32), colour=np.full(n, 0xFF2D1010, np.uint32), ids=ids,
                 min_bound=(0.0, 0.0, 0.0), max_bound=tuple(box), steps=2,
                 mc_samples=(32, 32, 32), seed=seed)


def _block_dims(n, aspect):
    """Integer lattice dims with product close to n and the given axis ratios."""
    a = np.asarray(aspect, np.float64)
    base = (n / np.prod(a)) ** (1.0 / 3.0)
    dims = np.maximum(2, np.rint(base * a)).astype(int)
    # adjust the longest axis so the count is as close to n as possible
    k = int(np.argmax(dims))
    rest = int(np.prod(dims)) // int(dims[k])
    dims[k] = max(2, int(round(n / rest)))
    return tuple(int(d) for d in dims)


def batch_scenes(n_scenes=64, lo=2000, hi=200000, seed=3, iterations=4, total=None,
                 spacing_range=(0.01, 0.05), box_factor=(2.0, 1.5, 1.5), mc_samples=(64, 64, 64)):
    """Independent dam breaks of uneven size for batched stepping.

    Particle counts N_s = exp(U(ln lo, ln hi)) (seeded), optionally rescaled so
    they sum to `total`; each scene is a jittered lattice block whose axes have
    random aspect factors U(0.5, 2) normalised to N_s, in the corner of its own
    box (block extent x box_factor: 2x in x, 1.5x in y and z), with its own
    world spacing (log-uniform in spacing_range) and so its own scale. Every
    scene uses the solver-unit recipe of the module docstring (h = 2 spacings,
    about 33 neighbours within h)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), n_scenes))
    if total is not None:
        counts = counts * (float(total) / counts.sum())
    out = []
    for s in range(n_scenes):
        aspect = rng.uniform(0.5, 2.0, 3)
        aspect = aspect / np.cbrt(np.prod(aspect))
        dims = _block_dims(int(round(counts[s])), aspect)
        sp = float(np.exp(rng.uniform(np.log(spacing_range[0]), np.log(spacing_range[1]))))
        sp = float(np.float32(sp))
        box = tuple(float(np.float32((d * sp) * f + 2 * sp)) for d, f in zip(dims, box_factor))
        sc = _scene("batch_%d_%d" % (s, int(np.prod(dims))), [(dims, (sp / 2,) * 3)], sp, box,
                    seed * 1000 + s, 10, mc_samples)
        sc.iterations = iterations
        out.append(sc)
    return out


def combined_scenes(seed=5):
    """32 scenes, N_s log-uniform in [50k, 500k] (seed 5) rescaled to a total
    of about 8,000,000 particles, 6 solver iterations, marching cubes at 512^3
    per scene."""
    return batch_scenes(32, 50000, 500000, seed=seed, iterations=6, total=8000000,
                        mc_samples=(512, 512, 512))


def icosphere(centre, radius, subdivisions=2):
    """Triangles (n, 3, 3) float32 of a sphere mesh: an icosahedron whose faces
    are split in four `subdivisions` times, vertices pushed onto the sphere;
    20 * 4**subdivisions triangles, outward winding. A collider mesh for
    tests and benchmarks."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.asarray(centre, np.float64)
    return V[np.array(faces)].astype(np.float32)
