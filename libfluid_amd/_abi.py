"""ctypes declarations of the C ABI in include/libfluid_amd.h.

The shared library is built in-tree (libfluid_amd/liblfamd.so, by
``__graft_entry__.build()`` or ``make -C libfluid_amd``). There is no CPU
fallback: loading fails loudly when the library is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblfamd.so")

LF_OK = 0
LF_ERR_ARG = -1
LF_ERR_HIP = -2
LF_ERR_MORTON = -3
LF_ERR_CAPACITY = -4
LF_ERR_STATE = -5
LF_ERR_NODEVICE = -6
LF_ERR_DIVERGED = -7

LF_FLUID = 0
LF_OBSTACLE = 1

LF_FLAG_DIFFUSE = 0x1
LF_FLAG_KAHAN = 0x2
LF_FLAG_FAST_MATH = 0x4
LF_FLAG_PACKED16 = 0x8

LF_MC_FIELD_REFERENCE = 0
LF_MC_FIELD_DENSITY = 1

LF_ARR_KEYS = 1
LF_ARR_ORDER = 2
LF_ARR_TABLE = 3
LF_ARR_PSTAR = 4
LF_ARR_LAMBDA = 5
LF_ARR_POSITION = 6
LF_ARR_VELOCITY = 7
LF_ARR_TYPE = 8
LF_ARR_DIFFUSED = 9
LF_ARR_FIELD = 10
LF_ARR_MCNORMAL = 11
LF_ARR_MCCOLOUR = 12
LF_ARR_CASES = 13
LF_ARR_DENSITY = 14
LF_ARR_LISTSTATS = 15
LF_ARR_CELLKEY = 16

_f3 = ctypes.c_float * 3
_u3 = ctypes.c_uint32 * 3


class PbfParams(ctypes.Structure):
    """lf_pbf_params: ClSphConfig (oclsph_type.h:6-13) + -DSPH_H + kernel constants."""
    _fields_ = [
        ("h", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("scale", ctypes.c_float),
        ("iterations", ctypes.c_uint32),
        ("gravity", _f3),
        ("min_bound", _f3),
        ("max_bound", _f3),
        ("rest_density", ctypes.c_float),
        ("cfm_epsilon", ctypes.c_float),
        ("corr_k", ctypes.c_float),
        ("corr_dq", ctypes.c_float),
        ("vd", ctypes.c_float),
        ("xsph_c", ctypes.c_float),
        ("flags", ctypes.c_uint32),
        ("density_tol", ctypes.c_float),
    ]


class McParams(ctypes.Structure):
    """lf_mc_params: ClMcConfig (oclsph_type.h:39-44) + the density-field lattice."""
    _fields_ = [
        ("field", ctypes.c_uint32),
        ("isolevel", ctypes.c_float),
        ("resolution", ctypes.c_float),
        ("particle_size", ctypes.c_float),
        ("particle_influence", ctypes.c_float),
        ("samples", _u3),
        ("attributes", ctypes.c_uint32),
    ]


class Well(ctypes.Structure):
    _fields_ = [("centre", _f3), ("force", ctypes.c_float)]


class Particle(ctypes.Structure):
    """lf_particle: the layout of fluid::Particle<size_t,float> (fluid.hpp:53-60)."""
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("type", ctypes.c_int32),
        ("mass", ctypes.c_float),
        ("position", _f3),
        ("velocity", _f3),
        ("colour", ctypes.c_uint32),
    ]


class MeshTriangle(ctypes.Structure):
    """lf_mesh_triangle: the layout of geometry::MeshTriangle<float> (geometry.hpp:41-44)."""
    _fields_ = [
        ("v0", _f3), ("v1", _f3), ("v2", _f3),
        ("n0", _f3), ("n1", _f3), ("n2", _f3),
        ("c", _u3),
    ]


class Source(ctypes.Structure):
    """lf_source: a particle source (fluid::Source<float>, fluid.hpp:147-159), without the tag."""
    _fields_ = [("centre", ctypes.c_float * 3), ("velocity", ctypes.c_float * 3),
                ("rate", ctypes.c_uint32), ("colour", ctypes.c_uint32)]


class Drain(ctypes.Structure):
    """lf_drain: a spherical drain (fluid::Drain<float>, fluid.hpp:162-169)."""
    _fields_ = [("centre", ctypes.c_float * 3), ("width", ctypes.c_float),
                ("depth", ctypes.c_float)]


class Diag(ctypes.Structure):
    """lf_diag: FP64 per-scene conserved quantities and density error."""
    _fields_ = [
        ("fluid", ctypes.c_uint64),
        ("mass", ctypes.c_double),
        ("momentum", ctypes.c_double * 3),
        ("density_error_mean", ctypes.c_double),
        ("density_error_max", ctypes.c_double),
        ("passes", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("pass_error", ctypes.c_double * 16),
    ]


class TuneResult(ctypes.Structure):
    """lf_tune_result: the configuration lf_autotune kept."""
    _fields_ = [
        ("threads", ctypes.c_uint32),
        ("persistent", ctypes.c_int32),
        ("skin", ctypes.c_float),
        ("ms_per_step", ctypes.c_float),
        ("trials", ctypes.c_uint32),
        ("particles_per_thread", ctypes.c_uint32),
        ("mc_tile", ctypes.c_uint32),
        ("mc_ms", ctypes.c_float),
    ]


class ExecParams(ctypes.Structure):
    """lf_exec_params: the execution configuration."""
    _fields_ = [
        ("threads", ctypes.c_uint32),
        ("persistent", ctypes.c_int32),
        ("skin", ctypes.c_float),
        ("particles_per_thread", ctypes.c_uint32),
        ("mc_tile", ctypes.c_uint32),
    ]


# numpy views of the same records
def particle_dtype():
    import numpy as np
    return np.dtype({
        "names": ["id", "type", "mass", "position", "velocity", "colour"],
        "formats": ["<u8", "<i4", "<f4", ("<f4", 3), ("<f4", 3), "<u4"],
        "offsets": [0, 8, 12, 16, 28, 40],
        "itemsize": 48,
    })


def triangle_dtype():
    import numpy as np
    return np.dtype({
        "names": ["v", "n", "c"],
        "formats": [("<f4", (3, 3)), ("<f4", (3, 3)), ("<u4", 3)],
        "offsets": [0, 36, 72],
        "itemsize": 84,
    })


_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I = ctypes.c_int
_U32 = ctypes.c_uint32

# name -> (restype, argtypes); every symbol include/libfluid_amd.h declares
SIGNATURES = {
    "lf_pbf_params_default": (None, [ctypes.POINTER(PbfParams), ctypes.c_float]),
    "lf_ctx_create": (_I, [_I, ctypes.POINTER(_P)]),
    "lf_ctx_destroy": (_I, [_P]),
    "lf_last_error": (ctypes.c_char_p, [_P]),
    "lf_sync": (_I, [_P]),
    "lf_set_guard": (_I, [_P, ctypes.c_uint32]),
    "lf_health": (_I, [_P, ctypes.POINTER(ctypes.c_uint32)]),
    "lf_state_upload": (_I, [_P, _SZ, _P, _P, _P, _P, _P]),
    "lf_state_download": (_I, [_P, _SZ, _P, _P, _P, _P, _P, _P]),
    "lf_state_size": (_I, [_P, ctypes.POINTER(_SZ)]),
    "lf_set_wells": (_I, [_P, _SZ, _P]),
    "lf_set_colliders": (_I, [_P, _SZ, _P]),
    "lf_wire_particles": (_I, [_P, _P, _SZ, ctypes.c_int64, ctypes.POINTER(_SZ)]),
    "lf_wire_triangles": (_I, [_P, _P, _SZ, ctypes.c_int64, ctypes.POINTER(_SZ)]),
    "lf_profile_only": (_I, [_P, ctypes.c_char_p]),
    "lf_set_sources": (_I, [_P, _SZ, _P]),
    "lf_set_drains": (_I, [_P, _SZ, _P]),
    "lf_set_queries": (_I, [_P, _SZ, _P]),
    "lf_query_results": (_I, [_P, _SZ, _P, _P]),
    "lf_grid_build": (_I, [_P, ctypes.POINTER(PbfParams)]),
    "lf_pbf_diffuse": (_I, [_P, ctypes.POINTER(PbfParams)]),
    "lf_pbf_lambda": (_I, [_P, ctypes.POINTER(PbfParams)]),
    "lf_pbf_delta": (_I, [_P, ctypes.POINTER(PbfParams)]),
    "lf_pbf_finalise": (_I, [_P, ctypes.POINTER(PbfParams)]),
    "lf_pbf_xsph": (_I, [_P, ctypes.POINTER(PbfParams)]),
    "lf_pbf_step": (_I, [_P, ctypes.POINTER(PbfParams), _U32]),
    "lf_mc_sample": (_I, [_P, _P, ctypes.POINTER(McParams)]),
    "lf_mc_classify_count": (_I, [_P, ctypes.POINTER(McParams)]),
    "lf_mc_emit": (_I, [_P, _P, ctypes.POINTER(McParams)]),
    "lf_mc_count": (_I, [_P, ctypes.POINTER(_SZ)]),
    "lf_mc_download": (_I, [_P, _P, _SZ, ctypes.POINTER(_SZ)]),
    "lf_mc_set_capacity": (_I, [_P, _SZ]),
    "lf_mc_mesh_static": (_I, [_P, _P, _P, _P, ctypes.c_float, ctypes.c_float, _P, _P, _SZ, _P]),
    "lf_advance": (_I, [_P, ctypes.POINTER(PbfParams), ctypes.POINTER(McParams), _P, _SZ, _P,
                        _SZ, ctypes.POINTER(_SZ), _P]),
    "lf_mc_field_upload": (_I, [_P, _P, _P, _P, ctypes.c_float, _P]),
    "lf_read": (_I, [_P, _I, _P, _SZ]),
    "lf_write": (_I, [_P, _I, _P, _SZ]),
    "lf_mc_keep_cases": (_I, [_P, _I]),
    "lf_grid_info": (_I, [_P, ctypes.POINTER(_U32), _P, _P]),
    "lf_mc_info": (_I, [_P, _P, _P, _P]),
    "lf_profile_enable": (_I, [_P, _I]),
    "lf_profile_reset": (_I, [_P]),
    "lf_kernel_stats": (_I, [_P, ctypes.c_char_p, _SZ]),
    "lf_copy_probe": (_I, [_P, _SZ, _I, ctypes.POINTER(ctypes.c_float)]),
    "lf_batch_upload": (_I, [_P, _U32, _P, _P, _P, _P, _P, _P, _P]),
    "lf_batch_step": (_I, [_P, _U32]),
    "lf_batch_info": (_I, [_P, ctypes.POINTER(_U32), _P]),
    "lf_batch_select": (_I, [_P, _U32]),
    "lf_exec_config": (_I, [_P, _U32, _I, ctypes.c_float]),
    "lf_exec_configure": (_I, [_P, ctypes.POINTER(ExecParams)]),
    "lf_autotune": (_I, [_P, _U32, ctypes.POINTER(TuneResult)]),
    "lf_graph_capture": (_I, [_P, _U32, _P]),
    "lf_graph_launch": (_I, [_P, _U32]),
    "lf_diagnostics": (_I, [_P, _U32, ctypes.POINTER(Diag)]),
}

_lib = None


def load():
    """Load liblfamd.so (raises when it is not built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libfluid_amd: %s is not built; run __graft_entry__.build() or "
            "make -C libfluid_amd" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
