"""ORACLE - test infrastructure, not product code.

ctypes binding of the CPU restatement of libfluid's PBF step and marching
cubes (oracle/pbf_oracle.hpp, built into oracle/liblforacle.so by
``make -C oracle``), and of the reference's own compiled headers in
oracle/_ref (``make -C oracle ref``, only where the reference tree exists).
Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module, and only as the checker.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(_HERE, "liblforacle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libref.so")

# stage ids of oc_stage
GRID, LAMBDA, DELTA, FINALISE, XSPH, DIFFUSE, COMMIT, STEP = range(8)
# extra array ids beyond LF_ARR_*
CALLER_POSITION, CALLER_VELOCITY, CALLER_MASS, CALLER_KEYS, CALLER_COLOUR = 101, 102, 103, 104, 105

_libs = {}


def lib(path=None):
    """The oracle library (ORACLE_LIB, or a build of it at `path`: the
    planted-mistake builds of tests/test_oracle_mutations.py)."""
    path = path or ORACLE_LIB
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.oc_create.restype = P
        L.oc_create.argtypes = [P, ctypes.c_int]
        L.oc_destroy.argtypes = [P]
        L.oc_load.argtypes = [P, ctypes.c_size_t, P, P, P, P, P]
        L.oc_set_wells.argtypes = [P, ctypes.c_size_t, P]
        L.oc_set_colliders.argtypes = [P, ctypes.c_size_t, P]
        L.oc_collide.restype = ctypes.c_int
        L.oc_collide.argtypes = [P, P, P]
        L.oc_stage.restype = ctypes.c_double
        L.oc_stage.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.oc_size.restype = ctypes.c_size_t
        L.oc_size.argtypes = [P, ctypes.c_int]
        L.oc_get_f64.argtypes = [P, ctypes.c_int, P]
        L.oc_set_f64.argtypes = [P, ctypes.c_int, P]
        L.oc_get_u32.argtypes = [P, ctypes.c_int, P]
        L.oc_get_i32_types.argtypes = [P, P]
        L.oc_grid_info.argtypes = [P, P, P, P]
        L.oc_consts.argtypes = [P, P]
        L.oc_w6.restype = ctypes.c_double
        L.oc_w6.argtypes = [P, ctypes.c_double]
        L.oc_grad_scale.restype = ctypes.c_double
        L.oc_grad_scale.argtypes = [P, ctypes.c_double]
        L.oc_morton.restype = ctypes.c_uint32
        L.oc_morton.argtypes = [ctypes.c_uint32] * 3
        L.oc_demorton.argtypes = [ctypes.c_uint32, P]
        L.oc_mc_ref_dims.argtypes = [P, ctypes.c_float, P]
        L.oc_mc_reference.argtypes = [P, P, P, P, P]
        L.oc_mc_density.argtypes = [P, P, P]
        L.oc_density_lattice.argtypes = [P, P, P, P, ctypes.c_int]
        L.oc_mc_cases.argtypes = [P, P, ctypes.c_double, ctypes.c_int, P]
        L.oc_mc_emit.restype = ctypes.c_size_t
        L.oc_mc_emit.argtypes = [P, P, P, P, ctypes.c_double, P, P, ctypes.c_double, ctypes.c_int,
                                 P, ctypes.c_size_t]
        L.oc_mc_table.restype = ctypes.c_int
        L.oc_mc_table.argtypes = [ctypes.c_int] * 3
        _libs[path] = L
    return _libs[path]


def ref_lib():
    """The reference's own curves.h / MC tables, compiled (None if absent)."""
    if not os.path.exists(REF_LIB):
        return None
    L = ctypes.CDLL(REF_LIB)
    sz = ctypes.c_size_t
    L.ref_morton.restype = sz
    L.ref_morton.argtypes = [sz, sz, sz]
    L.ref_decode.restype = sz
    L.ref_decode.argtypes = [sz, ctypes.c_int]
    L.ref_index3d.restype = sz
    L.ref_index3d.argtypes = [sz] * 6
    L.ref_to3d.argtypes = [sz, sz, sz, sz, ctypes.c_void_p]
    L.ref_mc_table.restype = ctypes.c_int
    L.ref_mc_table.argtypes = [ctypes.c_int] * 3
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class Pbf:
    """One oracle instance (FP32 reference semantics, or FP64)."""

    def __init__(self, params, fp64=False, lib_path=None):
        self._L = lib(lib_path)
        self.fp64 = fp64
        self._h = self._L.oc_create(ctypes.byref(params), 1 if fp64 else 0)
        self.params = params

    def __del__(self):
        try:
            self._L.oc_destroy(self._h)
        except Exception:
            pass

    def load(self, position, velocity, mass, ptype=None, colour=None):
        pos = np.ascontiguousarray(position, np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.ascontiguousarray(velocity, np.float64).reshape(n, 3)
        m = np.ascontiguousarray(np.broadcast_to(mass, (n,)), np.float64)
        t = np.zeros(n, np.int32) if ptype is None else np.ascontiguousarray(ptype, np.int32)
        c = (np.full(n, 0xFF2D1010, np.uint32) if colour is None
             else np.ascontiguousarray(colour, np.uint32))
        self._L.oc_load(self._h, n, _p(pos), _p(vel), _p(m), _p(t), _p(c))
        self.n = n

    def set_wells(self, wells):
        from libfluid_amd._abi import Well
        arr = (Well * max(1, len(wells)))()
        for i, w in enumerate(wells):
            arr[i].centre[:] = list(w.centre)
            arr[i].force = w.force
        self._L.oc_set_wells(self._h, len(wells), ctypes.cast(arr, ctypes.c_void_p))

    def set_colliders(self, triangles):
        t = np.ascontiguousarray(np.asarray(triangles, np.float32).reshape(-1, 9))
        self._L.oc_set_colliders(self._h, t.shape[0], _p(t))

    def collide(self, position, velocity):
        """collideTriangle2 on one point: (hit, position, velocity)."""
        p = np.array(position, np.float64)
        v = np.array(velocity, np.float64)
        hit = self._L.oc_collide(self._h, _p(p), _p(v))
        return bool(hit), p, v

    def stage(self, s, repeat=1):
        """Run one stage; returns wall seconds."""
        return self._L.oc_stage(self._h, s, repeat)

    def grid_build(self):
        return self.stage(GRID)

    def lambda_pass(self):
        return self.stage(LAMBDA)

    def delta_pass(self):
        return self.stage(DELTA)

    def finalise(self):
        return self.stage(FINALISE)

    def xsph(self):
        return self.stage(XSPH)

    def diffuse(self):
        return self.stage(DIFFUSE)

    def commit(self):
        return self.stage(COMMIT)

    def step(self, nsteps=1):
        return self.stage(STEP, nsteps)

    def get(self, which):
        from libfluid_amd import _abi as A
        size = self._L.oc_size(self._h, which)
        if which in (A.LF_ARR_KEYS, A.LF_ARR_ORDER, A.LF_ARR_TABLE, A.LF_ARR_DIFFUSED,
                     CALLER_KEYS, CALLER_COLOUR):
            out = np.empty(size, np.uint32)
            self._L.oc_get_u32(self._h, which, _p(out))
            return out
        out = np.empty(size, np.float64)
        self._L.oc_get_f64(self._h, which, _p(out))
        if which == A.LF_ARR_PSTAR:
            return out.reshape(-1, 4)
        if which in (A.LF_ARR_POSITION, A.LF_ARR_VELOCITY, CALLER_POSITION, CALLER_VELOCITY):
            return out.reshape(-1, 3)
        return out

    def set(self, which, data):
        data = np.ascontiguousarray(data, np.float64).ravel()
        self._L.oc_set_f64(self._h, which, _p(data))

    def types(self):
        out = np.empty(self.n, np.int32)
        self._L.oc_get_i32_types(self._h, _p(out))
        return out

    def grid_info(self):
        G = np.zeros(1, np.uint32)
        ext = np.zeros(3, np.uint32)
        mn = np.zeros(3, np.float64)
        self._L.oc_grid_info(self._h, _p(G), _p(ext), _p(mn))
        return int(G[0]), tuple(int(e) for e in ext), tuple(mn)

    def consts(self):
        out = np.zeros(4, np.float64)
        self._L.oc_consts(self._h, _p(out))
        return dict(poly6=out[0], spiky=out[1], wdq=out[2], hh=out[3])

    def w6(self, r):
        return self._L.oc_w6(self._h, r)

    def grad_scale(self, r):
        return self._L.oc_grad_scale(self._h, r)

    # marching cubes
    def mc_reference(self, mc):
        S = np.zeros(3, np.uint32)
        self._L.oc_mc_ref_dims(self._h, ctypes.c_float(mc.resolution), _p(S))
        total = int(np.prod(S.astype(np.int64)))
        field = np.empty(total, np.float64)
        normal = np.empty(4 * total, np.float64)
        col = np.empty(total, np.uint32)
        self._L.oc_mc_reference(self._h, ctypes.byref(mc), _p(field), _p(normal), _p(col))
        S = tuple(int(s) for s in S)
        return S, field.reshape(S), normal.reshape(S + (4,)), col.reshape(S)

    def mc_density(self, samples):
        S = np.asarray(samples, np.uint32)
        field = np.empty(int(np.prod(S.astype(np.int64))), np.float64)
        self._L.oc_mc_density(self._h, _p(S), _p(field))
        return field.reshape(tuple(int(s) for s in S))

    def density_lattice(self, samples):
        S = np.asarray(samples, np.uint32)
        o = np.zeros(3, np.float64)
        s = np.zeros(3, np.float64)
        self._L.oc_density_lattice(ctypes.byref(self.params), _p(S), _p(o), _p(s),
                                   1 if self.fp64 else 0)
        return o, s


def mc_cases(field, iso, fp64=False):
    field = np.ascontiguousarray(field, np.float64)
    S = np.asarray(field.shape, np.uint32)
    out = np.empty(tuple(int(s) - 1 for s in S), np.uint8)
    lib().oc_mc_cases(_p(S), _p(field), iso, 1 if fp64 else 0, _p(out))
    return out


def mc_emit(field, iso, origin, step, scale=1.0, normal=None, colour=None, fp64=False):
    """Triangles (lf_mesh_triangle records) in cube order."""
    from libfluid_amd._abi import triangle_dtype
    field = np.ascontiguousarray(field, np.float64)
    S = np.asarray(field.shape, np.uint32)
    o = np.asarray(origin, np.float64)
    s = np.asarray(step, np.float64)
    nrm = None if normal is None else np.ascontiguousarray(normal, np.float64)
    col = None if colour is None else np.ascontiguousarray(colour, np.uint32)
    L = lib()
    count = L.oc_mc_emit(_p(S), _p(field), _p(nrm) if nrm is not None else None,
                         _p(col) if col is not None else None, iso, _p(o), _p(s), scale,
                         1 if fp64 else 0, None, 0)
    out = np.zeros(max(1, count), dtype=triangle_dtype())
    L.oc_mc_emit(_p(S), _p(field), _p(nrm) if nrm is not None else None,
                 _p(col) if col is not None else None, iso, _p(o), _p(s), scale,
                 1 if fp64 else 0, _p(out), count)
    return out[:count]


def morton(x, y, z):
    return lib().oc_morton(x, y, z)


def demorton(k):
    out = np.zeros(3, np.uint32)
    lib().oc_demorton(k, _p(out))
    return tuple(int(v) for v in out)


def mc_table(which, ci, i=0):
    return lib().oc_mc_table(which, ci, i)
