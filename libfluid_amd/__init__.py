"""libfluid_amd: the MI355X (gfx950) Position Based Fluids step and its
marching-cubes surface extraction, behind libfluid's solver interface.

Python host side over the C ABI of include/libfluid_amd.h:

* ``Context``  - one device context: device-resident particle state and the
  stage-by-stage entry points (grid, lambda, delta, finalise, XSPH, diffuse,
  marching cubes), inspection and kernel timing.
* ``SphSolver`` - mirrors ``fluid::SphSolver<size_t,float>::advance``
  (include/fluid/fluid.hpp:222-232) as implemented by ``ocl::SphSolver``
  (include/fluid/oclsph.hpp:661-936): sources, drains, one step, marching
  cubes on the reference field, particles returned in Morton order.
* ``Config`` / ``Source`` / ``Drain`` / ``Query`` / ``WellSpec`` - the
  records of fluid.hpp:22-209.

The compute path is the HIP library only; importing works without a GPU but
every compute call needs liblfamd.so and a device, and fails loudly otherwise.
"""
import ctypes
import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

from . import _abi
from ._abi import (LF_ARR_CASES, LF_ARR_DENSITY, LF_ARR_DIFFUSED, LF_ARR_FIELD, LF_ARR_KEYS,
                   LF_ARR_LAMBDA, LF_ARR_MCCOLOUR, LF_ARR_MCNORMAL, LF_ARR_ORDER, LF_ARR_POSITION,
                   LF_ARR_PSTAR, LF_ARR_TABLE, LF_ARR_TYPE, LF_ARR_VELOCITY, LF_FLAG_DIFFUSE,
                   LF_FLAG_FAST_MATH, LF_FLAG_KAHAN, LF_FLAG_PACKED16, LF_FLUID, LF_MC_FIELD_DENSITY, LF_MC_FIELD_REFERENCE,
                   LF_OBSTACLE, McParams, PbfParams, particle_dtype, triangle_dtype)

__all__ = [
    "Context", "SphSolver", "Config", "Source", "Drain", "Query", "QueryResult", "WellSpec",
    "PbfParams", "McParams", "LfError", "pbf_params", "mc_reference_params",
    "mc_density_params", "particle_dtype", "triangle_dtype", "LF_FLUID", "LF_OBSTACLE",
    "LF_FLAG_DIFFUSE", "LF_FLAG_KAHAN", "LF_FLAG_FAST_MATH", "LF_FLAG_PACKED16",
]


class LfError(RuntimeError):
    """A non-zero status from the C ABI (text from lf_last_error)."""

    def __init__(self, code, message):
        super().__init__("libfluid_amd error %d: %s" % (code, message))
        self.code = code


def pbf_params(h, dt=1.0 / 120.0, scale=1.0, iterations=4, gravity=(0.0, -9.8, 0.0),
               min_bound=(0.0, 0.0, 0.0), max_bound=(1.0, 1.0, 1.0), xsph_c=0.0, flags=0,
               rest_density=6378.0, density_tol=0.0):
    """lf_pbf_params with the reference's kernel constants (oclsph_kernel.h:20-31)."""
    p = PbfParams()
    _abi.load().lf_pbf_params_default(ctypes.byref(p), ctypes.c_float(h))
    p.dt = dt
    p.scale = scale
    p.iterations = iterations
    p.gravity[:] = list(gravity)
    p.min_bound[:] = list(min_bound)
    p.max_bound[:] = list(max_bound)
    p.xsph_c = xsph_c
    p.flags = flags
    p.rest_density = rest_density
    p.density_tol = density_tol
    return p


def mc_reference_params(isolevel, resolution=1.0, attributes=1):
    """ClMcConfig as advance() sets it (oclsph.hpp:669-673)."""
    m = McParams()
    m.field = LF_MC_FIELD_REFERENCE
    m.isolevel = isolevel
    m.resolution = resolution
    m.particle_size = 25.0
    m.particle_influence = 0.5
    m.attributes = attributes
    return m


def mc_density_params(isolevel, samples, attributes=0):
    """Density-field marching cubes on a samples^3 lattice spanning the box."""
    m = McParams()
    m.field = LF_MC_FIELD_DENSITY
    m.isolevel = isolevel
    m.resolution = 1.0
    s = samples if isinstance(samples, (tuple, list)) else (samples,) * 3
    m.samples[:] = list(s)
    m.attributes = attributes
    return m


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class Context:
    """One lf_ctx: device-resident state on its own HIP stream."""

    def __init__(self, device=0):
        self._lib = _abi.load()
        h = ctypes.c_void_p()
        rc = self._lib.lf_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise LfError(rc, "lf_ctx_create(%d) failed (no HIP device?)" % device)
        self._h = h
        self.n = 0

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.lf_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, allow=()):
        if rc != 0 and rc not in allow:
            msg = self._lib.lf_last_error(self._h)
            raise LfError(rc, msg.decode() if msg else "")
        return rc

    # -- state ----------------------------------------------------------------
    def upload(self, position, velocity, mass, ptype=None, colour=None):
        position = np.ascontiguousarray(position, dtype=np.float32).reshape(-1, 3)
        n = position.shape[0]
        velocity = np.ascontiguousarray(velocity, dtype=np.float32).reshape(n, 3)
        mass = np.ascontiguousarray(np.broadcast_to(mass, (n,)), dtype=np.float32)
        ptype = (np.zeros(n, np.int32) if ptype is None
                 else np.ascontiguousarray(ptype, dtype=np.int32))
        colour = (np.full(n, 0xFF2D1010, np.uint32) if colour is None
                  else np.ascontiguousarray(colour, dtype=np.uint32))
        self._check(self._lib.lf_state_upload(self._h, n, _ptr(position), _ptr(velocity),
                                              _ptr(mass), _ptr(ptype), _ptr(colour)))
        self.n = n

    def download(self):
        n = self.n
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        mass = np.empty(n, np.float32)
        typ = np.empty(n, np.int32)
        col = np.empty(n, np.uint32)
        order = np.empty(n, np.uint32)
        self._check(self._lib.lf_state_download(self._h, n, _ptr(pos), _ptr(vel), _ptr(mass),
                                                _ptr(typ), _ptr(col), _ptr(order)))
        return dict(position=pos, velocity=vel, mass=mass, type=typ, colour=col, order=order)

    def set_wells(self, wells):
        arr = (_abi.Well * max(1, len(wells)))()
        for i, w in enumerate(wells):
            arr[i].centre[:] = list(w.centre)
            arr[i].force = w.force
        self._check(self._lib.lf_set_wells(self._h, len(wells), ctypes.cast(arr, ctypes.c_void_p)))

    def set_colliders(self, triangles):
        """Mesh colliders: float array (n, 3, 3) of triangle corners, world units."""
        t = np.ascontiguousarray(np.asarray(triangles, np.float32).reshape(-1, 9))
        self._check(self._lib.lf_set_colliders(self._h, t.shape[0], _ptr(t) if t.size else None))

    def set_sources(self, sources):
        """Sources appended on the device at the start of every lf_pbf_step:
        each an object with centre, velocity, rate and colour."""
        arr = (_abi.Source * max(1, len(sources)))()
        for i, src in enumerate(sources):
            arr[i].centre[:] = [float(v) for v in src.centre]
            arr[i].velocity[:] = [float(v) for v in src.velocity]
            arr[i].rate = int(src.rate)
            arr[i].colour = int(src.colour) & 0xFFFFFFFF
        self._check(self._lib.lf_set_sources(self._h, len(sources), arr if sources else None))

    def set_drains(self, drains):
        """Drains applied on the device after the sources of every lf_pbf_step."""
        arr = (_abi.Drain * max(1, len(drains)))()
        for i, d in enumerate(drains):
            arr[i].centre[:] = [float(v) for v in d.centre]
            arr[i].width = float(d.width)
            arr[i].depth = float(d.depth)
        self._check(self._lib.lf_set_drains(self._h, len(drains), arr if drains else None))

    def set_queries(self, points):
        """Query points (world units), answered by every lf_pbf_step after its grid build."""
        pts = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
        self._n_queries = pts.shape[0]
        self._check(self._lib.lf_set_queries(self._h, pts.shape[0], _ptr(pts) if pts.size else None))

    def query_results(self):
        """(counts, colours) of the last step's answers to the query points."""
        n = getattr(self, "_n_queries", 0)
        counts = np.zeros(n, np.uint32)
        colours = np.zeros(n, np.uint32)
        self._check(self._lib.lf_query_results(self._h, n, _ptr(counts), _ptr(colours)))
        return counts, colours

    def _wire(self, fn, buf, timestamp):
        need = ctypes.c_size_t()
        ts = ctypes.c_int64(int(timestamp))
        if buf is None:
            self._check(fn(self._h, None, 0, ts, ctypes.byref(need)))
            out = np.zeros(need.value, np.uint8)
            self._check(fn(self._h, _ptr(out), out.nbytes, ts, ctypes.byref(need)))
            return out.tobytes()
        arr = np.frombuffer(buf, np.uint8)  # a writable buffer: bytearray, mmap, numpy
        self._check(fn(self._h, arr.ctypes.data, arr.nbytes, ts, ctypes.byref(need)))
        return need.value

    def wire_particles(self, buf=None, timestamp=0):
        """The particle block of the mmap wire format, packed on the device.
        With buf (a writable buffer such as an mmap) it is written there and
        the block size returned; without, the block is returned as bytes."""
        return self._wire(self._lib.lf_wire_particles, buf, timestamp)

    def wire_triangles(self, buf=None, timestamp=0):
        """The meshTriangle block of the last marching-cubes emit."""
        return self._wire(self._lib.lf_wire_triangles, buf, timestamp)

    # -- stages ---------------------------------------------------------------
    def grid_build(self, p):
        self._check(self._lib.lf_grid_build(self._h, ctypes.byref(p)))

    def diffuse(self, p):
        self._check(self._lib.lf_pbf_diffuse(self._h, ctypes.byref(p)))

    def lambda_pass(self, p):
        self._check(self._lib.lf_pbf_lambda(self._h, ctypes.byref(p)))

    def delta_pass(self, p):
        self._check(self._lib.lf_pbf_delta(self._h, ctypes.byref(p)))

    def finalise(self, p):
        self._check(self._lib.lf_pbf_finalise(self._h, ctypes.byref(p)))

    def xsph(self, p):
        self._check(self._lib.lf_pbf_xsph(self._h, ctypes.byref(p)))

    def step(self, p, nsteps=1):
        self._check(self._lib.lf_pbf_step(self._h, ctypes.byref(p), nsteps))
        size = ctypes.c_size_t()
        self._check(self._lib.lf_state_size(self._h, ctypes.byref(size)))
        self.n = size.value  # sources and drains change the particle count

    def sync(self):
        self._check(self._lib.lf_sync(self._h))

    def set_guard(self, max_cell_particles):
        """Divergence guard: cells above this many particles (0: no limit) stop
        the pair passes; clears the health counters."""
        self._check(self._lib.lf_set_guard(self._h, int(max_cell_particles)))

    def health(self):
        """(non-finite p* count, overfull cells, largest overfull cell)."""
        out = (ctypes.c_uint32 * 3)()
        self._check(self._lib.lf_health(self._h, out))
        return tuple(int(v) for v in out)

    # -- batches, execution configuration, graphs ---------------------------
    def batch_upload(self, params, scenes):
        """Several scenes in one context. params: one PbfParams per scene;
        scenes: dicts with position, velocity, mass[, type, colour]."""
        counts = np.array([len(np.asarray(sc["position"]).reshape(-1, 3)) for sc in scenes],
                          dtype=np.uint64)
        n = int(counts.sum())
        pos = np.concatenate([np.asarray(sc["position"], np.float32).reshape(-1, 3)
                              for sc in scenes]) if n else np.zeros((0, 3), np.float32)
        vel = np.concatenate([np.asarray(sc["velocity"], np.float32).reshape(-1, 3)
                              for sc in scenes]) if n else np.zeros((0, 3), np.float32)
        mass = np.concatenate([np.broadcast_to(np.asarray(sc["mass"], np.float32),
                                               (int(c),)) for sc, c in zip(scenes, counts)])
        typ = np.concatenate([np.asarray(sc.get("type", np.zeros(int(c), np.int32)), np.int32)
                              for sc, c in zip(scenes, counts)])
        col = np.concatenate([np.asarray(sc.get("colour", np.full(int(c), 0xFF2D1010)),
                                         np.uint32) for sc, c in zip(scenes, counts)])
        arr = (_abi.PbfParams * len(params))(*params)
        pos, vel, mass, typ, col = (np.ascontiguousarray(a) for a in (pos, vel, mass, typ, col))
        self._check(self._lib.lf_batch_upload(
            self._h, len(params), ctypes.cast(arr, ctypes.c_void_p), _ptr(counts), _ptr(pos),
            _ptr(vel), _ptr(mass), _ptr(typ), _ptr(col)))
        self.n = n

    def batch_step(self, nsteps=1):
        self._check(self._lib.lf_batch_step(self._h, nsteps))

    def batch_info(self):
        """Scene offsets (n_scenes + 1 entries)."""
        S = ctypes.c_uint32()
        self._check(self._lib.lf_batch_info(self._h, ctypes.byref(S), None))
        off = np.zeros(S.value + 1, np.uint64)
        self._check(self._lib.lf_batch_info(self._h, ctypes.byref(S), _ptr(off)))
        return off.astype(np.int64)

    def batch_select(self, scene):
        self._check(self._lib.lf_batch_select(self._h, scene))

    def diagnostics(self, scene=0):
        """FP64 per-scene fluid count, mass, momentum and density error."""
        d = _abi.Diag()
        self._check(self._lib.lf_diagnostics(self._h, scene, ctypes.byref(d)))
        return dict(fluid=d.fluid, mass=d.mass, momentum=tuple(d.momentum),
                    density_error_mean=d.density_error_mean,
                    density_error_max=d.density_error_max, passes=d.passes,
                    pass_error=list(d.pass_error)[:max(0, d.passes)])

    def exec_config(self, threads=256, persistent=False, skin=0.15, particles_per_thread=1,
                    mc_tile=0):
        """Execution configuration (results do not change): workgroup size
        (128, 256, 512), the persistent work queue, the neighbour-list skin as
        a fraction of h, chunks per workgroup (1, 2, 4) and the density-field
        tile (0: 8x8x16 samples, 1: 4x16x16, 2: 8x8x8, 3: 8x8x4)."""
        e = _abi.ExecParams(threads, 1 if persistent else 0, skin, particles_per_thread, mc_tile)
        self._check(self._lib.lf_exec_configure(self._h, ctypes.byref(e)))

    def autotune(self, steps=3):
        r = _abi.TuneResult()
        self._check(self._lib.lf_autotune(self._h, steps, ctypes.byref(r)))
        return dict(threads=r.threads, persistent=bool(r.persistent), skin=r.skin,
                    particles_per_thread=r.particles_per_thread, mc_tile=r.mc_tile,
                    ms_per_step=r.ms_per_step, mc_ms=r.mc_ms, trials=r.trials)

    def graph_capture(self, nsteps=2, m=None):
        self._check(self._lib.lf_graph_capture(self._h, nsteps,
                                               ctypes.byref(m) if m is not None else None))

    def graph_launch(self, times=1):
        self._check(self._lib.lf_graph_launch(self._h, times))

    # -- marching cubes -----------------------------------------------------
    def mc_sample(self, p, m):
        """p None: the constants of the last step (a batch: the selected scene's)."""
        self._check(self._lib.lf_mc_sample(self._h, ctypes.byref(p) if p is not None else None,
                                           ctypes.byref(m)))

    def mc_classify_count(self, m):
        self._check(self._lib.lf_mc_classify_count(self._h, ctypes.byref(m)))

    def mc_emit(self, p, m):
        self._check(self._lib.lf_mc_emit(self._h, ctypes.byref(p) if p is not None else None,
                                         ctypes.byref(m)))

    def mc_field_upload(self, field, origin, step, scale=1.0):
        """Upload a scalar field (x-major, shape S) for classify / emit."""
        field = np.ascontiguousarray(field, np.float32)
        S = (ctypes.c_uint32 * 3)(*field.shape)
        o = (ctypes.c_float * 3)(*origin)
        st = (ctypes.c_float * 3)(*step)
        self._check(self._lib.lf_mc_field_upload(self._h, S, o, st, scale, _ptr(field)))

    def mesh_field(self, field, isolevel, origin, step, scale=1.0):
        """Marching cubes of an arbitrary scalar field on the GPU."""
        self.mc_field_upload(field, origin, step, scale)
        m = mc_density_params(isolevel, tuple(np.asarray(field).shape))
        self.mc_classify_count(m)
        self.mc_emit(None, m)
        return self.mc_triangles()

    def mesh_static(self, d_field, samples, origin, step, scale, isolevel, d_tri_pos, tri_cap,
                    d_ntri):
        """lf_mc_mesh_static: marching cubes of a field in device memory into
        device buffers (addresses as integers), allocating nothing; enqueued on
        the context's stream (sync() waits)."""
        S = (ctypes.c_uint32 * 3)(*samples)
        o = (ctypes.c_float * 3)(*origin)
        st = (ctypes.c_float * 3)(*step)
        self._check(self._lib.lf_mc_mesh_static(self._h, S, o, st, scale, isolevel, d_field,
                                                 d_tri_pos if tri_cap else None, tri_cap, d_ntri))

    def mc_count(self):
        n = ctypes.c_size_t()
        self._check(self._lib.lf_mc_count(self._h, ctypes.byref(n)))
        return n.value

    def mc_triangles(self, cap=None):
        total = self.mc_count()
        cap = total if cap is None else cap
        out = np.zeros(max(1, min(cap, total)), dtype=triangle_dtype())
        got = ctypes.c_size_t()
        self._check(self._lib.lf_mc_download(self._h, _ptr(out), min(cap, total),
                                             ctypes.byref(got)))
        return out[:min(cap, total)]

    def mc_set_capacity(self, max_triangles):
        self._check(self._lib.lf_mc_set_capacity(self._h, max_triangles))

    def mc_keep_cases(self, on=True):
        self._check(self._lib.lf_mc_keep_cases(self._h, 1 if on else 0))

    def mesh(self, p, m):
        """Sample, classify, emit; returns the triangles (synchronises)."""
        self.mc_sample(p, m)
        self.mc_classify_count(m)
        self.mc_emit(p, m)
        return self.mc_triangles()

    # -- inspection -----------------------------------------------------------
    def grid_info(self):
        G = ctypes.c_uint32()
        ext = (ctypes.c_uint32 * 3)()
        mn = (ctypes.c_float * 3)()
        self._check(self._lib.lf_grid_info(self._h, ctypes.byref(G), ext, mn))
        return G.value, tuple(ext), tuple(mn)

    def mc_info(self):
        S = (ctypes.c_uint32 * 3)()
        o = (ctypes.c_float * 3)()
        s = (ctypes.c_float * 3)()
        self._check(self._lib.lf_mc_info(self._h, S, o, s))
        return tuple(S), tuple(o), tuple(s)

    def read(self, array, count=None):
        n = self.n
        if array in (LF_ARR_KEYS, LF_ARR_ORDER, LF_ARR_DIFFUSED):
            out = np.empty(n, np.uint32)
        elif array == LF_ARR_TABLE:
            out = np.empty(self.grid_info()[0] if count is None else count, np.uint32)
        elif array == _abi.LF_ARR_CELLKEY:
            out = np.empty(n, np.uint32)
        elif array in (LF_ARR_PSTAR, LF_ARR_LAMBDA, LF_ARR_POSITION, LF_ARR_VELOCITY):
            out = np.empty((n, 4), np.float32)
        elif array == LF_ARR_TYPE:
            out = np.empty(n, np.uint8)
        elif array == LF_ARR_DENSITY:
            out = np.empty(n, np.float32)
        elif array == _abi.LF_ARR_LISTSTATS:
            out = np.empty(4, np.uint32)
        elif array == LF_ARR_FIELD:
            S = self.mc_info()[0]
            out = np.empty(S, np.float32)
        elif array == LF_ARR_MCNORMAL:
            S = self.mc_info()[0]
            out = np.empty(S + (4,), np.float32)
        elif array == LF_ARR_MCCOLOUR:
            S = self.mc_info()[0]
            out = np.empty(S, np.uint32)
        elif array == LF_ARR_CASES:
            S = self.mc_info()[0]
            out = np.empty(tuple(s - 1 for s in S), np.uint8)
        else:
            raise ValueError("unknown array %r" % array)
        self._check(self._lib.lf_read(self._h, array, _ptr(out), out.nbytes))
        return out

    def write(self, array, data):
        data = np.ascontiguousarray(data)
        self._check(self._lib.lf_write(self._h, array, _ptr(data), data.nbytes))

    # -- timing ---------------------------------------------------------------
    def profile(self, on=True):
        self._check(self._lib.lf_profile_enable(self._h, 1 if on else 0))

    def profile_only(self, name=None):
        """Time only the launches named `name` (None: every launch)."""
        self._check(self._lib.lf_profile_only(self._h, name.encode() if name else None))

    def profile_reset(self):
        self._check(self._lib.lf_profile_reset(self._h))

    def copy_probe(self, nbytes, reps=20):
        """Mean ms of one device-to-device copy of nbytes by a plain kernel."""
        ms = ctypes.c_float()
        self._check(self._lib.lf_copy_probe(self._h, nbytes, reps, ctypes.byref(ms)))
        return ms.value

    def kernel_stats(self):
        """{kernel name: (launches, total ms)} over the profiled launches."""
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(self._lib.lf_kernel_stats(self._h, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, count, ms = line.split()
            out[name] = (int(count), float(ms))
        return out


# ---- the drop-in solver -------------------------------------------------------

@dataclasses.dataclass
class WellSpec:
    """fluid::Well (fluid.hpp:137-144)."""
    centre: Sequence[float]
    force: float


@dataclasses.dataclass
class Source:
    """fluid::Source (fluid.hpp:146-159)."""
    centre: Sequence[float]
    velocity: Sequence[float]
    rate: int
    tag: int
    colour: int


@dataclasses.dataclass
class Drain:
    """fluid::Drain (fluid.hpp:161-169)."""
    centre: Sequence[float]
    width: float
    depth: float


@dataclasses.dataclass
class Query:
    """fluid::Query (fluid.hpp:22-30)."""
    id: int
    point: Sequence[float]


@dataclasses.dataclass
class QueryResult:
    """fluid::QueryResult (fluid.hpp:33-50)."""
    id: int
    point: Sequence[float]
    neighbours: int
    average_colour: int


@dataclasses.dataclass
class Config:
    """fluid::Config<float> (fluid.hpp:172-209)."""
    dt: float
    scale: float
    resolution: float
    isolevel: float
    iteration: int
    constant_force: Sequence[float]
    wells: List[WellSpec] = dataclasses.field(default_factory=list)
    sources: List[Source] = dataclasses.field(default_factory=list)
    drains: List[Drain] = dataclasses.field(default_factory=list)
    queries: List[Query] = dataclasses.field(default_factory=list)
    min_bound: Sequence[float] = (0.0, 0.0, 0.0)
    max_bound: Sequence[float] = (1.0, 1.0, 1.0)


@dataclasses.dataclass
class Result:
    """fluid::Result<float> (fluid.hpp:212-220): triangles as a numpy record
    array of lf_mesh_triangle (geometry::MeshTriangle<float> layout)."""
    triangles: np.ndarray
    queries: List[QueryResult]


class SphSolver:
    """ocl::SphSolver(h, kernelPath, device) (oclsph.hpp:170) as a HIP solver.

    ``advance(config, xs)`` follows ocl::SphSolver::advance
    (oclsph.hpp:661-936): sources add particles (:678-691), drains erase
    fluid particles (:693-706), an empty set returns an empty result
    (:711-715), one PBF step with colour diffusion runs on the device, the
    marching cubes run on the reference field, and ``xs`` (a numpy record
    array of ``particle_dtype()``) is returned permuted into this step's Morton
    order (:874-892). Returns ``(xs, Result)``: numpy arrays cannot grow in
    place, so the new particle array is returned rather than mutated.
    """

    def __init__(self, h, device=0, xsph_c=0.0):
        self.h = float(h)
        self.xsph_c = xsph_c
        self.ctx = Context(device)
        self._tri_cap = 1 << 20

    def _params(self, config):
        return pbf_params(self.h, dt=config.dt, scale=config.scale, iterations=config.iteration,
                          gravity=config.constant_force, min_bound=config.min_bound,
                          max_bound=config.max_bound, xsph_c=self.xsph_c, flags=LF_FLAG_DIFFUSE)

    def advance(self, config, xs):
        xs = _apply_sources_drains(config, xs, self.h)
        if len(xs) == 0:
            return xs, Result(np.zeros(0, dtype=triangle_dtype()), [])
        p = self._params(config)
        m = mc_reference_params(config.isolevel, config.resolution, attributes=1)
        lib = self.ctx._lib
        self.ctx.set_wells(config.wells)
        xs = np.ascontiguousarray(xs, dtype=particle_dtype())
        work = xs.copy()
        order = np.empty(len(work), np.uint32)
        ntri = ctypes.c_size_t()
        rc = lib.lf_advance(self.ctx._h, ctypes.byref(p), ctypes.byref(m), _ptr(work),
                            len(work), None, 0, ctypes.byref(ntri), _ptr(order))
        self.ctx._check(rc)
        self.ctx.n = len(work)
        tris = self.ctx.mc_triangles()
        queries = self._queries(config, work, xs["colour"][order]) if config.queries else []
        return work, Result(tris, queries)

    def _queries(self, config, xs, colours):
        """Spatial queries (oclsph.hpp:770-791): fluid particles of the query
        point's own cell, their count and mean colour (the colours before this
        step's diffusion, in this step's Morton order)."""
        G, _, min_ext = self.ctx.grid_info()
        table = self.ctx.read(LF_ARR_TABLE)
        out = []
        for q in config.queries:
            r = [np.float32(q.point[k]) / np.float32(config.scale) - np.float32(min_ext[k])
                 for k in range(3)]
            cells = [int(np.float32(r[k]) / np.float32(self.h)) if r[k] >= 0 else 0
                     for k in range(3)]
            z = _morton(*cells)
            if z < G and z + 1 < G:
                rng = slice(int(table[z]), int(table[z + 1]))
                fluid = xs["type"][rng] == LF_FLUID
                if fluid.any():
                    c = colours[rng][fluid].astype(np.uint64)
                    argb = np.stack([(c >> 24) & 255, (c >> 16) & 255, (c >> 8) & 255,
                                     c & 255], axis=1).astype(np.float32)
                    avg = argb.sum(axis=0) / np.float32(len(c))
                    packed = 0
                    for ch in avg:
                        packed = (packed << 8) | (int(ch) & 0xFF)
                    out.append(QueryResult(q.id, tuple(q.point), len(c), packed))
        return out


def _spread10(v):
    v &= 1023
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    v = (v | (v << 2)) & 0x09249249
    return v


def _morton(x, y, z):
    return _spread10(x) | (_spread10(y) << 1) | (_spread10(z) << 2)


def _apply_sources_drains(config, xs, h):
    """Sources then drains, as oclsph.hpp:678-706 (host side of advance)."""
    xs = np.asarray(xs, dtype=particle_dtype())
    added = []
    spacing = np.float32(h) * np.float32(config.scale) / np.float32(2)
    for src in config.sources:
        size = np.float32(math.sqrt(src.rate))
        width, depth = int(math.floor(size)), int(math.ceil(size))
        off = np.asarray(src.centre, np.float32) - (
            np.array([width, 0, depth], np.float32) / np.float32(2) * spacing)
        for x in range(width):
            for z in range(depth):
                rec = np.zeros(1, dtype=particle_dtype())
                rec["id"] = src.tag
                rec["type"] = LF_FLUID
                rec["mass"] = 1.0
                rec["colour"] = src.colour
                rec["position"] = off + np.array([x, 0, z], np.float32) * spacing
                rec["velocity"] = np.asarray(src.velocity, np.float32)
                added.append(rec)
    if added:
        xs = np.concatenate([xs] + added)
    if config.drains and len(xs):
        keep = np.ones(len(xs), bool)
        for d in config.drains:
            r = np.asarray(d.centre, np.float32) - xs["position"]
            dist = np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2])
            keep &= ~((xs["type"] != LF_OBSTACLE) & (dist < d.width))
        xs = xs[keep]
    return xs
