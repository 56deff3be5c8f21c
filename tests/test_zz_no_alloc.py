"""GPU: lf_mc_mesh_static, the marching-cubes entry point that allocates
nothing (device field in, device triangles and count out, block counts in a
fixed device array), against the allocating path of the same field
(lf_mc_field_upload + lf_mc_classify_count + lf_mc_emit): the same count and
the same vertex positions bit for bit, on two calls in a row, the second on a
smaller lattice (the shared array keeps the first call's counts), plus a cut
output (tri_cap below the count) and the size limit."""
import ctypes

import numpy as np
import pytest

import libfluid_amd as lf
from libfluid_amd._abi import LF_ERR_ARG

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _field(shape, seed):
    """A sum of Gaussian blobs minus a slope, with a plane of samples exactly at
    the isolevel 0.5: closed and open surfaces, ties on the comparison."""
    rng = np.random.default_rng(seed)
    axes = [np.linspace(0.0, 1.0, s, dtype=np.float64) for s in shape]
    x, y, z = np.meshgrid(*axes, indexing="ij")
    f = 0.3 * (1.0 - y)
    for _ in range(24):
        c = rng.uniform(0.05, 0.95, 3)
        r = rng.uniform(0.03, 0.12)
        f += np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (r * r))
    f = f.astype(np.float32)
    f[:, :, shape[2] // 2] = 0.5
    return np.ascontiguousarray(f)


def _static(c, field, origin, step, scale, iso, cap):
    d_field = torch.from_numpy(field.ravel()).to("cuda:0")
    d_tri = torch.full((max(cap, 1) * 9,), float("nan"), dtype=torch.float32, device="cuda:0")
    d_n = torch.full((1,), -1, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    print("lf_mc_mesh_static samples %s cap %d" % (field.shape, cap), flush=True)
    c.mesh_static(d_field.data_ptr(), field.shape, origin, step, scale, iso, d_tri.data_ptr(),
                  cap, d_n.data_ptr())
    c.sync()
    n = int(d_n.cpu().numpy().view(np.uint32)[0])
    tris = d_tri.cpu().numpy()[:min(n, cap) * 9].reshape(-1, 3, 3)
    return n, tris


CASES = (((97, 81, 73), 11, (0.0, -0.5, 0.25), (0.01, 0.0125, 0.02), 2.0),
         ((41, 48, 57), 12, (1.0, 0.0, -1.0), (0.025, 0.02, 0.0175), 0.5))


def test_static_mesh_matches_allocating_path_twice():
    with lf.Context(0) as c, lf.Context(0) as ref:
        for shape, seed, origin, step, scale in CASES:
            f = _field(shape, seed)
            want = ref.mesh_field(f, 0.5, origin, step, scale)
            n, got = _static(c, f, origin, step, scale, 0.5, len(want) + 7)
            assert n == len(want) and n > 1000, (shape, n, len(want))
            assert np.array_equal(got, want["v"]), shape
        # a cut output: the full count, the first tri_cap triangles
        n, got = _static(c, f, origin, step, scale, 0.5, 100)
        assert n == len(want) and got.shape == (100, 3, 3)
        assert np.array_equal(got, want["v"][:100])
        n, got = _static(c, f, origin, step, scale, 0.5, 0)
        assert n == len(want)


def test_static_mesh_size_limit():
    with lf.Context(0) as c:
        dummy = torch.zeros(16, dtype=torch.float32, device="cuda:0")
        S = (ctypes.c_uint32 * 3)(1025, 1025, 258)  # 2^28 + 2^20 cubes
        o = (ctypes.c_float * 3)(0, 0, 0)
        st = (ctypes.c_float * 3)(1, 1, 1)
        rc = c._lib.lf_mc_mesh_static(c._h, S, o, st, 1.0, 0.5, dummy.data_ptr(),
                                      dummy.data_ptr(), 1, dummy.data_ptr())
        assert rc == LF_ERR_ARG
        assert "at most 268435456" in c._lib.lf_last_error(c._h).decode()
