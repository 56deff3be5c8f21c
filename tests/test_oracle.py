"""CPU tests that pin the oracle (oracle/, the CPU restatement of libfluid's
PBF step and marching cubes) before it is trusted as the GPU checker.

Each check is chosen so a plausible mistake fails it: the reference's own
curves.h and marching-cubes tables compiled from the reference tree
(oracle/_ref), the reference tables' SHA-256 digests, closed forms (lattice
density, kernel normalisation, finalise), invariants (sort, cell table,
candidate sets, symmetric delta, XSPH on a uniform field) and a sphere mesh.
"""
import hashlib
import math
import struct

import numpy as np
import pytest

import libfluid_amd as lf
import oracle
from libfluid_amd import scenes
from libfluid_amd._abi import (LF_ARR_DENSITY, LF_ARR_KEYS, LF_ARR_LAMBDA, LF_ARR_ORDER,
                               LF_ARR_POSITION, LF_ARR_PSTAR, LF_ARR_TABLE, LF_ARR_VELOCITY)

REF = oracle.ref_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (reference tree absent)")


def interleave_bits(x, y, z):
    """Morton code by definition: bit i of x, y, z -> bits 3i, 3i+1, 3i+2."""
    k = 0
    for i in range(10):
        k |= ((x >> i) & 1) << (3 * i)
        k |= ((y >> i) & 1) << (3 * i + 1)
        k |= ((z >> i) & 1) << (3 * i + 2)
    return k


# ---- Morton codes ---------------------------------------------------------------

def test_morton_matches_bit_definition():
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1024, (2000, 3)):
        assert oracle.morton(int(x), int(y), int(z)) == interleave_bits(int(x), int(y), int(z))
    for x, y, z in [(0, 0, 0), (1023, 1023, 1023), (1, 0, 0), (0, 1, 0), (0, 0, 1), (512, 3, 77)]:
        k = oracle.morton(x, y, z)
        assert k == interleave_bits(x, y, z)
        assert oracle.demorton(k) == (x, y, z)


@needs_ref
def test_morton_matches_reference_curves_h():
    """The reference's own curves.h (zCurveGridIndexAtCoord / coordAtZCurveGridIndex*)."""
    rng = np.random.default_rng(1)
    pts = list(rng.integers(0, 1024, (5000, 3))) + [(x, y, z) for x in range(8) for y in range(8)
                                                    for z in range(8)]
    for x, y, z in pts:
        k = REF.ref_morton(int(x), int(y), int(z))
        assert k == oracle.morton(int(x), int(y), int(z))
        assert (REF.ref_decode(k, 0), REF.ref_decode(k, 1), REF.ref_decode(k, 2)) == (x, y, z)
    # the reference wraps x-1 at 0 through size_t arithmetic: 10-bit wrap
    assert REF.ref_morton(2**64 - 1, 5, 5) == oracle.morton(1023, 5, 5)
    assert REF.ref_morton(1024, 5, 5) == oracle.morton(0, 5, 5)


@needs_ref
def test_index3d_to3d_match_reference():
    import ctypes
    out = (ctypes.c_size_t * 3)()
    for (x, y, z), dims in [((1, 2, 3), (4, 5, 6)), ((63, 0, 62), (64, 64, 64)), ((0, 0, 0), (2, 2, 2))]:
        i = REF.ref_index3d(x, y, z, *dims)
        assert i == (x * dims[1] + y) * dims[2] + z
        REF.ref_to3d(i, *dims, out)
        assert tuple(out) == (x, y, z)


# ---- marching-cubes tables ----------------------------------------------------------

EDGE_SHA = "2572f4b459bce46d930e57b2c261ef8143b6d43157272c698c64d70dd295b596"
TRI_SHA = "19bf7699e214903d72c94c296546f2e31337d637a1e4b118c3108a0f428e809b"
NVERT_SHA = "6108ce43ea7db94a0bafc9accecb54f2585c46edc02bf45a968c3024dd5c79aa"


def _tables(get):
    edge = [get(0, ci, 0) for ci in range(256)]
    tri = [[get(1, ci, i) for i in range(16)] for ci in range(256)]
    nv = [get(2, ci, 0) for ci in range(256)]
    return edge, tri, nv


def test_mc_tables_match_reference_digests():
    edge, tri, nv = _tables(oracle.mc_table)
    assert hashlib.sha256(struct.pack("<256H", *edge)).hexdigest() == EDGE_SHA
    assert hashlib.sha256(b"".join(struct.pack("16b", *r) for r in tri)).hexdigest() == TRI_SHA
    assert hashlib.sha256(bytes(nv)).hexdigest() == NVERT_SHA
    assert sum(n // 3 for n in nv) == 820 and max(n // 3 for n in nv) == 5


@needs_ref
def test_mc_tables_match_reference_build():
    edge, tri, nv = _tables(oracle.mc_table)
    redge, rtri, rnv = _tables(REF.ref_mc_table)
    assert edge == redge and tri == rtri and nv == rnv
    # the device copy in mc.h (255 sentinel) carries the same data
    for ci in range(256):
        assert REF.ref_mc_table(10, ci, 0) == edge[ci]
        assert REF.ref_mc_table(12, ci, 0) == nv[ci]
        for i in range(16):
            v = REF.ref_mc_table(11, ci, i)
            assert (v if v != 255 else -1) == tri[ci][i]


def test_edge_table_closed_form():
    corners = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5),
               (2, 6), (3, 7)]
    for ci in range(256):
        bits = sum(1 << e for e, (a, b) in enumerate(corners) if ((ci >> a) & 1) != ((ci >> b) & 1))
        assert oracle.mc_table(0, ci) == bits
        # triangles only use cut edges
        row = [oracle.mc_table(1, ci, i) for i in range(16)]
        for e in row:
            if e == -1:
                break
            assert bits & (1 << e)


# ---- grid geometry ------------------------------------------------------------------

def test_extent_kats():
    o8 = oracle.Pbf(scenes.dam_break_8k().params())
    o8.load(np.zeros((1, 3)), np.zeros((1, 3)), np.ones(1))
    o8.grid_build()
    assert o8.grid_info()[:2] == (8112, (24, 14, 14))
    o1 = oracle.Pbf(scenes.dam_break_1m().params())
    o1.load(np.zeros((1, 3)), np.zeros((1, 3)), np.ones(1))
    o1.grid_build()
    assert o1.grid_info()[:2] == (3727478, (164, 83, 83))
    # truncation KAT: 2.1f / 0.025f truncates to 83 in FP32
    assert int(np.float32(2.1) / np.float32(0.025)) == 83


@pytest.fixture(scope="module")
def o8k():
    sc = scenes.dam_break_8k()
    o = oracle.Pbf(sc.params())
    o.load(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
    o.grid_build()
    return sc, o


def test_sort_is_stable_permutation(o8k):
    sc, o = o8k
    keys, order = o.get(LF_ARR_KEYS), o.get(LF_ARR_ORDER)
    caller_keys = o.get(oracle.CALLER_KEYS)
    assert np.array_equal(np.sort(order), np.arange(sc.n))
    G = o.grid_info()[0]
    assert np.array_equal(keys, np.minimum(caller_keys[order], G))
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    same = np.diff(keys.astype(np.int64)) == 0
    assert np.all(np.diff(order.astype(np.int64))[same] > 0)


def test_keys_match_float32_formula(o8k):
    """key = Morton(trunc((p* - minExt)/h)) with p* = fma(fma(m g, dt, v), dt, x/scale)."""
    sc, o = o8k
    G, ext, mn = o.grid_info()
    p = sc.params()
    x = sc.position.astype(np.float32)
    g = np.float32(-9.8) * sc.mass.astype(np.float32)
    dt = np.float32(p.dt)
    vy = (g.astype(np.float64) * dt).astype(np.float32)  # fma of exact product: v0 = 0
    pstar = x.copy()
    pstar[:, 1] = (vy.astype(np.float64) * dt + x[:, 1].astype(np.float64)).astype(np.float32)
    q = ((pstar - np.asarray(mn, np.float32)) / np.float32(p.h)).astype(np.float32)
    cells = np.floor(q).astype(np.int64)
    want = [interleave_bits(*c) for c in cells]
    assert np.array_equal(o.get(oracle.CALLER_KEYS), np.asarray(want, np.uint32))


def test_cell_table_is_lower_bound(o8k):
    sc, o = o8k
    keys, table = o.get(LF_ARR_KEYS), o.get(LF_ARR_TABLE)
    G = table.shape[0]
    assert np.array_equal(table, np.searchsorted(keys, np.arange(G), side="left"))


def test_candidate_sets_brute_force():
    """27-cell candidates = all particles whose cell differs by at most 1 per
    axis (brute force over all pairs), excluding the always-empty last cell."""
    sc = scenes.small_block(n_side=7)
    o = oracle.Pbf(sc.params())
    o.load(sc.position, sc.velocity, sc.mass)
    o.grid_build()
    keys, table = o.get(LF_ARR_KEYS), o.get(LF_ARR_TABLE)
    G = len(table)
    cells = np.array([oracle.demorton(int(k)) for k in keys])
    for a in range(0, len(keys), 17):
        cand = []
        cx, cy, cz = cells[a]
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    code = oracle.morton((cx + dx) & 1023, (cy + dy) & 1023, (cz + dz) & 1023)
                    if code + 1 < G:
                        cand += list(range(table[code], table[code + 1]))
        brute = [b for b in range(len(keys)) if np.all(np.abs(cells[b] - cells[a]) <= 1)
                 and keys[b] + 1 < G]
        assert sorted(cand) == brute


# ---- SPH kernels and the constraint passes -------------------------------------------

def test_poly6_normalised_and_spiky_closed_form():
    p = lf.pbf_params(0.125)
    o = oracle.Pbf(p, fp64=True)
    h = 0.125
    r = np.linspace(0.0, h, 20001)
    w = np.array([o.w6(x) for x in r])
    integral = np.trapezoid(4 * math.pi * r * r * w, r)
    assert abs(integral - 1.0) < 1e-6
    assert o.w6(h * 1.0001) == 0.0
    # spiky gradient magnitude: 45/(pi h^6) (h - r)^2
    for rr in (0.01, 0.05, 0.1):
        assert math.isclose(-o.grad_scale(rr) * rr, 45 / (math.pi * h ** 6) * (h - rr) ** 2,
                            rel_tol=1e-12)
    assert o.grad_scale(0.0) == 0.0 and o.grad_scale(h * 1.01) == 0.0


def _lattice64(n_side, s, origin):
    ii = np.arange(n_side, dtype=np.float64)
    g = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    return origin + g * s


def test_lambda_lattice_kat_exact():
    """Interior particle of an unjittered lattice, s = 1/16, h = 1/8 (exact in
    binary): rho = poly6 * sum over the 33 points within h of (h^2 - r^2)^3."""
    s, h = 1 / 16, 1 / 8
    pos = _lattice64(12, s, 0.5)
    p = lf.pbf_params(h, gravity=(0, 0, 0), max_bound=(2, 2, 2))
    o = oracle.Pbf(p, fp64=True)
    o.load(pos, np.zeros_like(pos), np.ones(len(pos)))
    o.grid_build()
    o.lambda_pass()
    rho, lam = o.get(LF_ARR_DENSITY), o.get(LF_ARR_LAMBDA)
    poly6 = 315 / (64 * math.pi * h ** 9)
    ks = [0] + [1] * 6 + [2] * 12 + [3] * 8 + [4] * 6
    want = poly6 * sum((h * h - k * s * s) ** 3 for k in ks)
    centre = np.argmin(np.abs(o.get(LF_ARR_PSTAR)[:, :3] - (0.5 + 6 * s)).sum(1))
    assert math.isclose(rho[centre], want, rel_tol=1e-12)
    assert math.isclose(lam[centre], -(want / 6378 - 1) / 600, rel_tol=1e-9)


def test_lambda_lattice_kat_half_h_spacing():
    """s = 0.05, h = 0.1, m = 1: rho = 8078.2013, lambda = -4.4429e-4 (float32
    positions, hence the 2e-6 tolerance)."""
    pos = _lattice64(12, 0.05, 0.3)
    p = lf.pbf_params(0.1, gravity=(0, 0, 0), max_bound=(2, 1, 1))
    o = oracle.Pbf(p, fp64=True)
    o.load(pos, np.zeros_like(pos), np.ones(len(pos)))
    o.grid_build()
    o.lambda_pass()
    rho = o.get(LF_ARR_DENSITY)
    i = int(np.argmax(rho))
    assert math.isclose(rho[i], 8078.2013, rel_tol=2e-6)
    assert math.isclose(o.get(LF_ARR_LAMBDA)[i], -4.4429e-4, rel_tol=1e-4)


def test_delta_zero_for_symmetric_particle():
    s, h = 1 / 16, 1 / 8
    pos = _lattice64(12, s, 0.5)
    p = lf.pbf_params(h, gravity=(0, 0, 0), max_bound=(2, 2, 2))
    o = oracle.Pbf(p, fp64=True)
    o.load(pos, np.zeros_like(pos), np.ones(len(pos)))
    o.grid_build()
    before = o.get(LF_ARR_PSTAR)[:, :3].copy()
    o.lambda_pass()
    o.delta_pass()
    after = o.get(LF_ARR_PSTAR)[:, :3]
    interior = np.all((before > 0.5 + 2.5 * s) & (before < 0.5 + 8.5 * s), axis=1)
    assert interior.sum() > 100
    assert np.abs(after[interior] - before[interior]).max() < 1e-15
    # boundary particles do move
    assert np.abs(after[~interior] - before[~interior]).max() > 1e-6


def test_finalise_closed_form_isolated_particle():
    """One isolated fluid particle: only gravity acts, so
    v_out = ((v0 + g dt) dt / dt + v0) * 0.49 and x_out = x + (v0 + g dt) dt."""
    p = lf.pbf_params(0.1, gravity=(0, -9.8, 0), max_bound=(1, 1, 1))
    o = oracle.Pbf(p, fp64=True)
    x0 = np.array([[0.5, 0.5, 0.5]])
    v0 = np.array([[0.25, 0.5, -0.125]])
    o.load(x0, v0, np.ones(1))
    o.step()
    dt = np.float64(np.float32(p.dt))
    g = np.float64(np.float32(-9.8))
    vp = v0 + np.array([[0, g * dt, 0]])
    want_x = x0 + vp * dt
    want_v = ((want_x - x0) / dt + v0) * np.float64(np.float32(0.49))
    assert np.allclose(o.get(oracle.CALLER_POSITION), want_x, rtol=0, atol=1e-14)
    assert np.allclose(o.get(oracle.CALLER_VELOCITY), want_v, rtol=1e-12, atol=1e-14)


def test_xsph_zero_for_uniform_velocity():
    sc = scenes.small_block(n_side=6)
    p = sc.params(xsph_c=0.01)
    p.gravity[:] = [0, 0, 0]
    p.iterations = 0
    o = oracle.Pbf(p, fp64=True)
    v = np.tile([[0.3, -0.2, 0.1]], (sc.n, 1))
    o.load(sc.position, v, sc.mass)
    o.grid_build()
    o.finalise()
    before = o.get(LF_ARR_VELOCITY).copy()
    o.xsph()
    assert np.abs(o.get(LF_ARR_VELOCITY) - before).max() < 1e-12


def test_fp32_oracle_close_to_fp64():
    sc = scenes.dam_break_8k()
    p = sc.params()
    o32, o64 = oracle.Pbf(p), oracle.Pbf(p, fp64=True)
    for o in (o32, o64):
        o.load(sc.position, sc.velocity, sc.mass)
        o.grid_build()
        o.lambda_pass()
    assert np.array_equal(o32.get(LF_ARR_KEYS), o64.get(LF_ARR_KEYS))
    r32, r64 = o32.get(LF_ARR_DENSITY), o64.get(LF_ARR_DENSITY)
    assert np.max(np.abs(r32 - r64) / r64) < 1e-5


def test_density_field_matches_fp64_brute_force():
    """The benchmark density field rho(x) = sum_b m_b 315/(64 pi h^9) (h^2 - d^2)^3
    over |x - p_b| < h, from an FP64 sum over every particle (no cells, KD-tree
    radius search) on a non-cubic lattice: pins the lattice geometry and axis
    order, the poly6 constant, the cube, the masses and the 27-cell candidate
    set of both oracle precisions."""
    from scipy.spatial import cKDTree
    sc = scenes.dam_break_8k()
    p = sc.params()
    S = (48, 24, 20)
    h = float(np.float32(scenes.SOLVER_H))  # the parameter block holds h as FP32
    poly6 = 315.0 / (64.0 * math.pi * h ** 9)
    for fp64, rel in ((True, 1e-11), (False, 1e-5)):
        o = oracle.Pbf(p, fp64=fp64)
        o.load(sc.position, sc.velocity, sc.mass)
        o.grid_build()
        field = o.mc_density(S)
        ps = o.get(LF_ARR_PSTAR)
        assert np.array_equal(ps[:, 3], np.ones(sc.n))
        origin, step = o.density_lattice(S)
        ax = [origin[k] + np.arange(S[k]) * step[k] for k in range(3)]
        X = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
        want = np.zeros(len(X))
        tree = cKDTree(ps[:, :3])
        for i, nb in enumerate(tree.query_ball_point(X, h)):
            if nb:
                d2 = ((ps[nb, :3] - X[i]) ** 2).sum(1)
                want[i] = (ps[nb, 3] * poly6 * np.maximum(h * h - d2, 0.0) ** 3).sum()
        want = want.reshape(S)
        assert (want > 0).sum() > 2000
        assert np.max(np.abs(field - want)) <= rel * want.max(), fp64


# ---- marching cubes ---------------------------------------------------------------

def test_mc_cases_are_thresholds():
    rng = np.random.default_rng(3)
    f = rng.uniform(0, 1, (5, 6, 7))
    ci = oracle.mc_cases(f, 0.5)
    off = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    for x in range(4):
        for y in range(5):
            for z in range(6):
                want = sum(1 << i for i, (a, b, c) in enumerate(off) if f[x + a, y + b, z + c] < 0.5)
                assert ci[x, y, z] == want


def test_mc_sphere_area_and_closure():
    S = 48
    r0 = 0.3
    ax = np.linspace(0, 1, S)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    f = r0 - np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)  # > 0 inside
    step = 1.0 / (S - 1)
    tris = oracle.mc_emit(f, 0.0, (0, 0, 0), (step,) * 3, 1.0, fp64=True)
    v = tris["v"].astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1).sum()
    assert abs(area - 4 * math.pi * r0 ** 2) / (4 * math.pi * r0 ** 2) < 0.01
    # vertices on the sphere
    rad = np.linalg.norm(v.reshape(-1, 3) - 0.5, axis=1)
    assert np.abs(rad - r0).max() < step
    # watertight: every edge shared by exactly two triangles
    q = np.round(v.reshape(-1, 3) / step * 4096).astype(np.int64)
    _, inv = np.unique(q, axis=0, return_inverse=True)
    inv = inv.reshape(-1, 3)
    edges = np.sort(np.concatenate([inv[:, [0, 1]], inv[:, [1, 2]], inv[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)


def test_collider_closed_form():
    """collideTriangle2 (oclsph_collision.h:25-58): a point above the triangle
    (0,0,0),(1,0,0),(0,0,1) projects onto its plane and the velocity reflects;
    a point whose projection falls outside, or whose plane is farther than 40
    units, is untouched; the first hit in order wins."""
    sc = scenes.small_block(n_side=4)
    for fp64 in (False, True):
        o = oracle.Pbf(sc.params(), fp64=fp64)
        o.set_colliders(np.array([[(0, 0, 0), (1, 0, 0), (0, 0, 1)],
                                  [(0, 0.25, 0), (1, 0.25, 0), (0, 0.25, 1)]], np.float32))
        hit, p, v = o.collide((0.2, 0.5, 0.2), (0.3, -1.0, 0.1))
        # the second triangle's plane (y = 0.25) comes second: the first wins
        assert hit and np.allclose(p, (0.2, 0.0, 0.2), atol=1e-7)
        assert np.allclose(v, (0.3, 1.0, 0.1), atol=1e-7)  # reflected about y
        hit, p, v = o.collide((0.8, 0.5, 0.8), (0.0, -1.0, 0.0))  # u + v > 1: outside both
        assert not hit and np.allclose(p, (0.8, 0.5, 0.8)) and np.allclose(v, (0, -1, 0))
        hit, _, _ = o.collide((0.2, 41.0, 0.2), (0, 0, 0))  # plane 41 units away
        assert not hit


def test_host_sources_and_drains_closed_form():
    """Sources (oclsph.hpp:678-691): floor(sqrt(rate)) x ceil(sqrt(rate))
    particles on an (x, z) lattice of spacing h*scale/2 centred on the source,
    appended after the existing particles; drains (oclsph.hpp:693-706) erase
    fluid inside the sphere of radius width and keep obstacles."""
    from libfluid_amd._abi import particle_dtype
    import types
    h, scale = 0.1, 0.5
    spacing = np.float32(h) * np.float32(scale) / np.float32(2)
    src = lf.Source(centre=(1.0, 2.0, 3.0), velocity=(0.5, 0.0, -1.0), rate=18, tag=3,
                    colour=0x11223344)
    xs = np.zeros(3, dtype=particle_dtype())
    xs["position"] = [(0.0, 0.0, 0.0), (5.0, 5.0, 5.0), (0.05, 0.0, 0.0)]
    xs["type"] = [lf.LF_FLUID, lf.LF_FLUID, lf.LF_OBSTACLE]
    cfg = types.SimpleNamespace(sources=[src], drains=[], scale=scale)
    out = lf._apply_sources_drains(cfg, xs, h)
    assert len(out) == 3 + 4 * 5
    new = out[3:]
    assert (new["type"] == lf.LF_FLUID).all() and (new["mass"] == 1).all()
    assert (new["colour"] == 0x11223344).all() and (new["id"] == 3).all()
    # x outer, z inner; width 4 in x, depth 5 in z; y on the centre
    xsr = np.round((new["position"][:, 0] - 1.0) / spacing + 2.0).astype(int)
    zsr = np.round((new["position"][:, 2] - 3.0) / spacing + 2.5).astype(int)
    assert xsr.tolist() == [x for x in range(4) for _ in range(5)]
    assert zsr.tolist() == [z for _ in range(4) for z in range(5)]
    assert (new["position"][:, 1] == np.float32(2.0)).all()
    assert np.allclose(new["position"][:, 0].mean(), 1.0 - spacing / 2, atol=1e-6)
    # the lattice's centre of mass sits half a spacing below the centre in x
    # (width 4 is even) and in z (offset depth/2 = 2.5 spacings, 5 points)
    assert np.allclose(new["position"][:, 2].mean(), 3.0 - spacing / 2, atol=1e-6)
    drain = lf.Drain(centre=(0.0, 0.0, 0.0), width=0.1, depth=0.0)
    cfg = types.SimpleNamespace(sources=[], drains=[drain], scale=scale)
    out = lf._apply_sources_drains(cfg, xs, h)
    # the fluid particle at the centre goes, the far one and the obstacle stay
    assert out["position"].tolist() == [[5.0, 5.0, 5.0], [np.float32(0.05), 0.0, 0.0]]
    # a distance exactly equal to width is kept (strict <)
    drain = lf.Drain(centre=(5.0, 5.0, 4.5), width=0.5, depth=0.0)
    cfg = types.SimpleNamespace(sources=[], drains=[drain], scale=scale)
    assert len(lf._apply_sources_drains(cfg, xs, h)) == 3


def test_diffuse_brute_force_over_cells():
    """sph_diffuse (oclsph_kernel.h:281-303) restated independently of the
    oracle's cell walk: every non-obstacle particle whose cell is within one
    cell of a's on each axis (the 27 cells, no distance test) adds its colour
    and counts; the mean times 1.33 is mixed into a's colour with weight dt/750,
    clamped to [8, 255] and converted toward zero per channel; obstacles keep
    their colour. dt = 75 makes the weight 0.1, so a dropped neighbour, a wrong
    count or a missing 1.33 moves channels by whole units."""
    from libfluid_amd._abi import LF_ARR_DIFFUSED, LF_OBSTACLE
    rng = np.random.default_rng(11)
    n, h = 600, 0.1
    pos = rng.uniform(0.35, 0.65, (n, 3))  # interior cells only: no wrap of the 27
    colour = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    ptype = np.where(rng.random(n) < 0.15, LF_OBSTACLE, 0).astype(np.int32)
    p = lf.pbf_params(h, dt=75.0, gravity=(0, 0, 0), max_bound=(1, 1, 1))
    o = oracle.Pbf(p, fp64=True)
    o.load(pos, np.zeros_like(pos), np.ones(n), ptype, colour)
    o.grid_build()
    o.diffuse()
    order = o.get(LF_ARR_ORDER).astype(np.int64)
    ps = o.get(LF_ARR_PSTAR)[:, :3]
    assert np.array_equal(ps, pos[order])  # zero velocity and gravity: p* = x, sorted
    _, _, mn = o.grid_info()
    cell = np.floor((ps - np.asarray(mn)) / h).astype(np.int64)
    cs = colour[order].view(np.uint8).reshape(n, 4).astype(np.float64)
    ts = ptype[order]
    got = o.get(LF_ARR_DIFFUSED).astype(np.uint32).view(np.uint8).reshape(n, 4)
    a = 75.0 / 750.0
    near_integer = 0
    for s in range(n):
        if ts[s] == LF_OBSTACLE:
            assert np.array_equal(got[s], cs[s].astype(np.uint8))
            continue
        nb = (np.abs(cell - cell[s]).max(1) <= 1) & (ts != LF_OBSTACLE)
        y = cs[nb].sum(0) / nb.sum() * 1.33
        out = np.clip(cs[s] + (y - cs[s]) * a, 8.0, 255.0)
        want = np.trunc(out)
        edge = np.abs(out - np.round(out)) < 1e-9  # double rounding may land either side
        near_integer += int(edge.sum())
        assert np.array_equal(got[s][~edge], want[~edge].astype(np.uint8)), s
        assert np.all(np.abs(got[s][edge].astype(int) - np.round(out[edge])) <= 1)
    assert near_integer < n // 10


def test_mc_reference_attributes_brute_force():
    """mc_lattice's value, normal and colour (oclsph_kernel.h:434-441, 497-529)
    restated as a brute-force sum over every particle, without cells: sample
    a = (minExt + (x, y, z) * h / res) * scale; each non-obstacle particle b
    with |p_b - a| < h * scale adds size / |l|^infl to the value,
    -infl * size * l / |l|^infl to the normal (l = p_b - a, normalised at the
    end) and its colour to a sum whose mean, converted toward zero, is the
    sample's colour. Interior samples only: the reference's clamped cells
    repeat edge cells there."""
    from libfluid_amd._abi import LF_ARR_DIFFUSED, LF_OBSTACLE
    rng = np.random.default_rng(12)
    n, h = 600, 0.1
    pos = rng.uniform(0.3, 0.7, (n, 3))
    colour = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    ptype = np.where(rng.random(n) < 0.15, LF_OBSTACLE, 0).astype(np.int32)
    p = lf.pbf_params(h, gravity=(0, 0, 0), max_bound=(1, 1, 1))
    o = oracle.Pbf(p, fp64=True)
    o.load(pos, np.zeros_like(pos), np.ones(n), ptype, colour)
    o.grid_build()
    m = lf.mc_reference_params(isolevel=100.0, resolution=1.0)
    S, field, normal, col = o.mc_reference(m)
    order = o.get(LF_ARR_ORDER).astype(np.int64)
    cs = o.get(LF_ARR_DIFFUSED).astype(np.uint32)
    assert np.array_equal(cs, colour[order])  # no diffusion pass ran
    xs, ts = pos[order], ptype[order]
    cb = cs.view(np.uint8).reshape(n, 4).astype(np.float64)
    _, ext, mn = o.grid_info()
    size, infl = float(m.particle_size), float(m.particle_influence)
    hf = float(np.float32(h))  # SPH_H is a float constant: the lattice step and threshold use it
    checked = filled = 0
    for x in range(1, min(S[0], ext[0] - 1)):
        for y in range(1, min(S[1], ext[1] - 1)):
            for z in range(1, min(S[2], ext[2] - 1)):
                a = np.asarray(mn) + np.array([x, y, z], np.float64) * hf
                l = xs - a
                ln = np.sqrt((l * l).sum(1))
                sel = (ln < hf) & (ts != LF_OBSTACLE)
                den = ln[sel] ** infl
                v = (size / den).sum()
                nrm = (-infl * size * l[sel] / den[:, None]).sum(0)
                if sel.any():
                    nrm /= np.sqrt((nrm * nrm).sum())
                    c = np.trunc(cb[sel].sum(0) / sel.sum()).astype(np.uint8)
                    filled += 1
                else:
                    c = np.zeros(4, np.uint8)
                assert math.isclose(field[x, y, z], v, rel_tol=1e-12, abs_tol=1e-300)
                assert np.allclose(normal[x, y, z, :3], nrm, rtol=0, atol=1e-12)
                assert np.array_equal(np.array([col[x, y, z]], np.uint32).view(np.uint8), c)
                checked += 1
    assert checked > 1000 and filled > 100
