"""GPU parity: the HIP path through the C ABI against the CPU restatement
(oracle/) on identical seeded inputs.

The FP32 path follows one arithmetic contract on both sides (explicit FMAs,
correctly rounded sqrt/div, the reference's 27-cell order, stable sort), so
the bar is bit-exact for keys, cell tables, orders, lambda, density, p*,
positions, velocities, colours, MC fields, case indices and triangles.
Comparisons use == on values (so -0 == +0).
"""
import numpy as np
import pytest

import libfluid_amd as lf
import oracle
from libfluid_amd import scenes
from libfluid_amd._abi import (LF_ARR_LISTSTATS, LF_ARR_CASES, LF_ARR_DENSITY, LF_ARR_DIFFUSED, LF_ARR_FIELD,
                               LF_ARR_KEYS, LF_ARR_LAMBDA, LF_ARR_MCCOLOUR, LF_ARR_MCNORMAL,
                               LF_ARR_ORDER, LF_ARR_POSITION, LF_ARR_PSTAR, LF_ARR_TABLE,
                               LF_ARR_VELOCITY, LF_ERR_MORTON)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = lf.Context(0)
    yield c
    c.close()


def _pair(ctx, sc, p, fp64=False):
    ctx.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
    o = oracle.Pbf(p, fp64=fp64)
    o.load(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
    return o


def _eq(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b).astype(a.dtype) if np.asarray(b).dtype != a.dtype else np.asarray(b)
    if a.shape != b.shape:
        raise AssertionError("%s: shape %s vs %s" % (what, a.shape, b.shape))
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else (a == b))
    if bad.any():
        idx = np.argwhere(bad)[:5]
        raise AssertionError("%s: %d of %d differ, first %s: gpu %s oracle %s" % (
            what, int(bad.sum()), bad.size, idx.tolist(), a[tuple(idx[0])], b[tuple(idx[0])]))


SCENES = {
    "8k": scenes.dam_break_8k,
    "small": lambda: scenes.small_block(n_side=9),
}


@pytest.mark.parametrize("name", ["8k", "small"])
def test_grid_bit_exact(ctx, name):
    sc = SCENES[name]()
    p = sc.params()
    o = _pair(ctx, sc, p)
    ctx.grid_build(p)
    o.grid_build()
    G, ext, mn = ctx.grid_info()
    oG, oext, omn = o.grid_info()
    assert (G, ext) == (oG, oext)
    _eq(ctx.read(LF_ARR_KEYS), o.get(LF_ARR_KEYS), "sorted keys")
    _eq(ctx.read(LF_ARR_ORDER), o.get(LF_ARR_ORDER), "sort order")
    _eq(ctx.read(LF_ARR_TABLE), o.get(LF_ARR_TABLE), "cell table")
    _eq(ctx.read(LF_ARR_PSTAR).astype(np.float64), o.get(LF_ARR_PSTAR), "p*")


def test_grid_1m_bit_exact(ctx):
    sc = scenes.dam_break_1m()
    p = sc.params()
    o = _pair(ctx, sc, p)
    ctx.grid_build(p)
    o.grid_build()
    G, ext, _ = ctx.grid_info()
    assert G == 3727478 and ext == (164, 83, 83)
    _eq(ctx.read(LF_ARR_KEYS), o.get(LF_ARR_KEYS), "sorted keys 1M")
    _eq(ctx.read(LF_ARR_ORDER), o.get(LF_ARR_ORDER), "sort order 1M")
    _eq(ctx.read(LF_ARR_TABLE), o.get(LF_ARR_TABLE), "cell table 1M")


@pytest.mark.parametrize("name", ["8k", "small"])
@pytest.mark.parametrize("kahan", [False, True])
def test_constraint_passes_bit_exact(ctx, name, kahan):
    sc = SCENES[name]()
    p = sc.params(flags=lf.LF_FLAG_KAHAN if kahan else 0)
    o = _pair(ctx, sc, p)
    ctx.grid_build(p)
    o.grid_build()
    for it in range(p.iterations):
        ctx.lambda_pass(p)
        o.lambda_pass()
        _eq(ctx.read(LF_ARR_LAMBDA)[:, 3].astype(np.float64), o.get(LF_ARR_LAMBDA),
            "lambda it%d" % it)
        _eq(ctx.read(LF_ARR_DENSITY).astype(np.float64), o.get(LF_ARR_DENSITY), "density it%d" % it)
        ctx.delta_pass(p)
        o.delta_pass()
        _eq(ctx.read(LF_ARR_PSTAR)[:, :3].astype(np.float64), o.get(LF_ARR_PSTAR)[:, :3],
            "p* it%d" % it)
    ctx.finalise(p)
    o.finalise()
    _eq(ctx.read(LF_ARR_POSITION)[:, :3].astype(np.float64), o.get(LF_ARR_POSITION), "x")
    _eq(ctx.read(LF_ARR_VELOCITY)[:, :3].astype(np.float64), o.get(LF_ARR_VELOCITY), "v")
    ctx.xsph(p)
    o.xsph()
    _eq(ctx.read(LF_ARR_VELOCITY)[:, :3].astype(np.float64), o.get(LF_ARR_VELOCITY), "v xsph")


def test_diffuse_bit_exact(ctx):
    sc = scenes.dam_break_8k()
    rng = np.random.default_rng(5)
    sc.colour = rng.integers(0, 2**32, sc.n, dtype=np.uint64).astype(np.uint32)
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    o = _pair(ctx, sc, p)
    ctx.grid_build(p)
    o.grid_build()
    ctx.diffuse(p)
    o.diffuse()
    _eq(ctx.read(LF_ARR_DIFFUSED), o.get(LF_ARR_DIFFUSED), "diffused colour")


@pytest.mark.parametrize("steps", [1, 10])
def test_steps_bit_exact_8k(ctx, steps):
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    o = _pair(ctx, sc, p)
    ctx.step(p, steps)
    o.step(steps)
    g = ctx.download()
    _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "position")
    _eq(g["velocity"].astype(np.float64), o.get(oracle.CALLER_VELOCITY), "velocity")
    _eq(g["colour"], o.get(oracle.CALLER_COLOUR), "colour")


def test_steps_bit_exact_small_h(ctx):
    """A solver h below 1e-4: the lambda and delta passes take the compiler's
    full IEEE sqrt and division (their shortened forms are exact only for
    1e-4 <= h <= 1e4). The 8k scene in solver units shrunk by f = 9e-4 (h and
    the world scale), with gravity, rest density, cfm and the s_corr constant
    scaled so the solver sees the same motion up to rounding (gravity ~ f,
    rho0 ~ f^-3, cfm ~ f^-2, corr_k ~ f^2)."""
    import dataclasses
    sc = scenes.dam_break_8k()
    f = 9e-4
    sc = dataclasses.replace(sc, h=sc.h * f, scale=sc.scale / f,
                             gravity=tuple(g * f for g in sc.gravity))
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    assert p.h < 1e-4
    p.rest_density = p.rest_density / f ** 3
    p.cfm_epsilon = p.cfm_epsilon / f ** 2
    p.corr_k = p.corr_k * f ** 2
    o = _pair(ctx, sc, p)
    ctx.grid_build(p)
    o.grid_build()
    for it in range(p.iterations):
        ctx.lambda_pass(p)
        o.lambda_pass()
        lam = ctx.read(LF_ARR_LAMBDA)[:, 3]
        _eq(lam.astype(np.float64), o.get(LF_ARR_LAMBDA), "lambda it%d" % it)
        assert np.abs(lam).max() > 0.0
        ctx.delta_pass(p)
        o.delta_pass()
        _eq(ctx.read(LF_ARR_PSTAR)[:, :3].astype(np.float64), o.get(LF_ARR_PSTAR)[:, :3],
            "p* it%d" % it)
    ctx.finalise(p)
    o.finalise()
    _eq(ctx.read(LF_ARR_POSITION)[:, :3].astype(np.float64), o.get(LF_ARR_POSITION), "x")
    ctx.xsph(p)
    o.xsph()
    _eq(ctx.read(LF_ARR_VELOCITY)[:, :3].astype(np.float64), o.get(LF_ARR_VELOCITY), "v xsph")
    # whole steps from the start
    o = _pair(ctx, sc, p)
    ctx.step(p, 3)
    o.step(3)
    g = ctx.download()
    _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "position")
    _eq(g["velocity"].astype(np.float64), o.get(oracle.CALLER_VELOCITY), "velocity")


def test_step_1m_bit_exact(ctx):
    sc = scenes.dam_break_1m()
    p = sc.params()
    o = _pair(ctx, sc, p)
    ctx.step(p, 1)
    o.step(1)
    g = ctx.download()
    _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "position 1M")
    _eq(g["velocity"].astype(np.float64), o.get(oracle.CALLER_VELOCITY), "velocity 1M")
    # the 256^3 density field of the stepped state (the staged-tile path with
    # its out-of-reach cells cut, and the global-memory path for crowded tiles)
    m = sc.mc()
    ctx.mc_sample(p, m)
    field = ctx.read(LF_ARR_FIELD)
    ofield = o.mc_density(m.samples)
    assert (field > 0).sum() > 1_000_000
    _eq(field.astype(np.float64).reshape(ofield.shape), ofield, "density field 1M 256^3")


def test_steps_1m_collapse_bit_exact(ctx):
    """The first 12 steps of the 1M dam break, where the constraint passes move
    particles farther than the neighbour lists' skin and many walks fall back
    to the 27 cells: bit for bit against the oracle's steps."""
    sc = scenes.dam_break_1m()
    p = sc.params()
    o = _pair(ctx, sc, p)
    fell = 0
    for _ in range(12):
        ctx.step(p, 1)
        fell += int(ctx.read(LF_ARR_LISTSTATS)[1:].astype(np.int64).sum())
    o.step(12)
    g = ctx.download()
    _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "position 1M, 12 steps")
    _eq(g["velocity"].astype(np.float64), o.get(oracle.CALLER_VELOCITY), "velocity 1M, 12 steps")
    assert fell > 10000, "the walks did not fall back to the cells (%d)" % fell


def test_far_neighbours_fall_back_bit_exact(ctx):
    """List entries are 24-bit offsets from the particle's own sorted position;
    a particle with a candidate 2^23 or more positions away (possible only in a
    context of more than 2^23 + 1 particles) walks its cells instead. A 272^3
    block, 20,123,648 particles, where neighbours across the coarsest Morton
    planes lie ~10M positions apart: one step (one iteration) matches the
    oracle bit for bit, and some lists did fall back."""
    s = scenes.SOLVER_SPACING
    sc = scenes._scene("block_20m", [((272, 272, 272), (s / 2,) * 3)], s, (27.2, 14.0, 14.0),
                       11, 1, (8, 8, 8))
    p = sc.params(iterations=1)
    o = _pair(ctx, sc, p)
    ctx.step(p, 1)
    stats = ctx.read(LF_ARR_LISTSTATS)
    o.step(1)
    g = ctx.download()
    _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "position 20M")
    _eq(g["velocity"].astype(np.float64), o.get(oracle.CALLER_VELOCITY), "velocity 20M")
    assert stats[1] > 0, "no lambda-pass lane walked cells"
    del o


def test_steps_with_marching_cubes_between_bit_exact(ctx):
    """Marching cubes between steps, as a renderer runs them: they run on a
    second stream of the context beside the next step, which writes the other
    of two alternating p* / cell-table buffers. Four steps issued without a
    host read between them, then new wells and two more: the state and the
    last density field equal the oracle's bit for bit."""
    sc = scenes.dam_break_8k()
    p = sc.params()
    o = _pair(ctx, sc, p)
    m = sc.mc()

    def step_and_mesh():
        ctx.step(p, 1)
        ctx.mc_sample(p, m)
        ctx.mc_classify_count(m)
        ctx.mc_emit(p, m)
        o.step(1)

    for _ in range(4):
        step_and_mesh()
    g = ctx.download()
    _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "position after 4")
    wells = [lf.WellSpec(centre=(1.0, 0.5, 0.5), force=40.0)]
    ctx.set_wells(wells)
    o.set_wells(wells)
    try:
        for _ in range(2):
            step_and_mesh()
    finally:
        ctx.set_wells([])  # the context is shared by the module's tests
    g = ctx.download()
    _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "position after 6")
    _eq(g["velocity"].astype(np.float64), o.get(oracle.CALLER_VELOCITY), "velocity after 6")
    field = ctx.read(LF_ARR_FIELD)
    ofield = o.mc_density(m.samples)
    _eq(field.astype(np.float64).reshape(ofield.shape), ofield, "field after 6")


def test_density_mc_bit_exact(ctx):
    sc = scenes.dam_break_8k()
    p = sc.params()
    o = _pair(ctx, sc, p)
    ctx.step(p, 1)
    o.step(1)
    # the oracle's sorted state after its step equals the GPU's (checked above)
    m = sc.mc()
    ctx.mc_keep_cases(True)
    ctx.mc_sample(p, m)
    ctx.mc_classify_count(m)
    ctx.mc_emit(p, m)
    field = ctx.read(LF_ARR_FIELD)
    ofield = o.mc_density(m.samples)
    _eq(field.astype(np.float64), ofield, "density field")
    cases = ctx.read(LF_ARR_CASES)
    _eq(cases, oracle.mc_cases(field.astype(np.float64), m.isolevel), "case indices")
    tris = ctx.mc_triangles()
    origin, step = o.density_lattice(m.samples)
    otris = oracle.mc_emit(field.astype(np.float64), m.isolevel, origin, step, 1.0)
    assert len(tris) == len(otris) > 0
    _eq(tris["v"], otris["v"], "triangle vertices")


@pytest.mark.parametrize("case", ["masses", "obstacles", "written"])
def test_density_field_other_masses_bit_exact(ctx, case):
    """The density field when the particles do not share one mass (its
    16-byte staging records), when obstacles are present (their p* record
    carries mass 0), and after positions were written with lf_write (which
    may change masses, so the field no longer assumes one)."""
    sc = scenes.dam_break_8k()
    p = sc.params()
    mass, typ, pos = sc.mass.copy(), sc.type.copy(), sc.position.copy()
    if case == "masses":
        mass = (1.0 + 0.25 * (np.arange(sc.n) % 4)).astype(np.float32)
    elif case == "obstacles":
        typ[::97] = lf.LF_OBSTACLE
    ctx.upload(pos, sc.velocity, mass, typ, sc.colour)
    o = oracle.Pbf(p)
    o.load(pos, sc.velocity, mass, typ, sc.colour)
    ctx.step(p, 1)
    o.step(1)
    if case == "written":
        x = ctx.read(LF_ARR_POSITION)
        ctx.write(LF_ARR_POSITION, x)  # the same bytes back
    m = sc.mc()
    ctx.mc_sample(p, m)
    _eq(ctx.read(LF_ARR_FIELD).astype(np.float64), o.mc_density(m.samples), "density field")


def test_reference_mc_bit_exact(ctx):
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    o = _pair(ctx, sc, p)
    ctx.step(p, 1)
    o.step(1)
    m = lf.mc_reference_params(isolevel=100.0, resolution=1.0)
    ctx.mc_sample(p, m)
    S, ofield, onormal, ocol = o.mc_reference(m)
    assert ctx.mc_info()[0] == S
    _eq(ctx.read(LF_ARR_FIELD).astype(np.float64), ofield, "metaball field")
    _eq(ctx.read(LF_ARR_MCNORMAL).astype(np.float64), onormal, "field normals")
    _eq(ctx.read(LF_ARR_MCCOLOUR), ocol, "field colours")
    ctx.mc_classify_count(m)
    ctx.mc_emit(p, m)
    tris = ctx.mc_triangles()
    G, ext, mn = o.grid_info()
    step = np.float32(p.h) / np.float32(m.resolution)
    otris = oracle.mc_emit(ofield, m.isolevel, mn, (step,) * 3, 1.0, normal=onormal, colour=ocol)
    assert len(tris) == len(otris) > 0
    _eq(tris["v"], otris["v"], "triangle vertices")
    _eq(tris["n"], otris["n"], "triangle normals")
    _eq(tris["c"], otris["c"], "triangle colours")


def test_advance_drop_in(ctx):
    """lf_advance through the SphSolver mirror: Morton-permuted particles and
    the reference-field mesh, against the oracle's step + mesh."""
    sc = scenes.dam_break_8k()
    solver = lf.SphSolver(sc.h)
    xs = np.zeros(sc.n, dtype=lf.particle_dtype())
    xs["id"] = sc.ids + 1000
    xs["type"] = sc.type
    xs["mass"] = sc.mass
    xs["position"] = sc.position
    xs["velocity"] = sc.velocity
    xs["colour"] = sc.colour
    cfg = lf.Config(dt=sc.dt, scale=1.0, resolution=1.0, isolevel=100.0, iteration=4,
                    constant_force=sc.gravity, min_bound=sc.min_bound, max_bound=sc.max_bound)
    out, res = solver.advance(cfg, xs)
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE, xsph_c=0.0)
    o = oracle.Pbf(p)
    o.load(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
    o.step()
    order = o.get(LF_ARR_ORDER)
    _eq(out["id"], xs["id"][order], "ids in Morton order")
    _eq(out["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "positions")
    _eq(out["velocity"].astype(np.float64), o.get(oracle.CALLER_VELOCITY), "velocities")
    _eq(out["colour"], o.get(oracle.CALLER_COLOUR), "colours")
    m = lf.mc_reference_params(isolevel=100.0, resolution=1.0)
    S, ofield, onormal, ocol = o.mc_reference(m)
    G, ext, mn = o.grid_info()
    step = np.float32(sc.h)
    otris = oracle.mc_emit(ofield, 100.0, mn, (step,) * 3, 1.0, normal=onormal, colour=ocol)
    assert len(res.triangles) == len(otris)
    _eq(res.triangles["v"], otris["v"], "mesh vertices")
    solver.ctx.close()


def test_empty_and_obstacles(ctx):
    sc = scenes.small_block(n_side=6)
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    # all obstacles: nothing moves
    t = np.ones(sc.n, np.int32)
    ctx.upload(sc.position, sc.velocity, sc.mass, t, sc.colour)
    ctx.step(p, 2)
    g = ctx.download()
    assert np.array_equal(np.sort(g["position"], axis=0), np.sort(sc.position, axis=0))
    # mixed fluid/obstacle, bit-exact against the oracle
    t = (np.arange(sc.n) % 3 == 0).astype(np.int32)
    ctx.upload(sc.position, sc.velocity, sc.mass, t, sc.colour)
    o = oracle.Pbf(p)
    o.load(sc.position, sc.velocity, sc.mass, t, sc.colour)
    ctx.step(p, 2)
    o.step(2)
    g = ctx.download()
    _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "mixed positions")
    # empty input
    ctx.upload(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
    ctx.step(p, 1)
    assert ctx.download()["position"].shape == (0, 3)


def test_outside_bounds_and_coincident(ctx):
    """Particles far outside the padded extent and exactly coincident pairs."""
    sc = scenes.small_block(n_side=6)
    pos = sc.position.copy()
    pos[0] = (-3.0, 0.5, 0.5)      # far below minExt: cell coordinate 0
    pos[1] = (5.0, 5.0, 5.0)       # beyond maxExt: key >= G
    pos[3] = pos[2]                # coincident pair
    p = sc.params()
    ctx.upload(pos, sc.velocity, sc.mass, sc.type, sc.colour)
    o = oracle.Pbf(p)
    o.load(pos, sc.velocity, sc.mass, sc.type, sc.colour)
    ctx.step(p, 1)
    o.step(1)
    g = ctx.download()
    _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "positions")
    assert np.isfinite(g["position"]).all()


def test_morton_extent_error(ctx):
    sc = scenes.small_block(n_side=4)
    p = sc.params()
    p.max_bound[0] = 200.0  # 2000+ cells on x: beyond the 10-bit Morton axis
    ctx.upload(sc.position, sc.velocity, sc.mass)
    with pytest.raises(lf.LfError) as e:
        ctx.grid_build(p)
    assert e.value.code == LF_ERR_MORTON


@pytest.mark.parametrize("h", [0.01, 0.1])
def test_sort_properties_random_keys(ctx, h):
    """Size-independent properties on a large random cloud: permutation,
    sortedness, stability, table = lower bound. h = 0.01: ~2 particles per
    cell; h = 0.1: ~700 per cell, every bucket ordered by the large-bucket sort."""
    rng = np.random.default_rng(11)
    n = 2_000_000
    pos = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    p = lf.pbf_params(h, gravity=(0, 0, 0), max_bound=(1, 1, 1))
    ctx.set_guard(0)
    ctx.upload(pos, np.zeros_like(pos), np.ones(n, np.float32))
    ctx.grid_build(p)
    keys = ctx.read(LF_ARR_KEYS)
    order = ctx.read(LF_ARR_ORDER)
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(order), np.arange(n, dtype=np.uint32))
    same = np.diff(keys.astype(np.int64)) == 0
    assert np.all(np.diff(order.astype(np.int64))[same] > 0), "sort is not stable"
    table = ctx.read(LF_ARR_TABLE)
    G = table.shape[0]
    assert np.array_equal(table, np.searchsorted(keys, np.arange(G, dtype=np.uint32),
                                                 side="left").astype(np.uint32))
    ctx.set_guard(256)


def test_1m_hundred_steps_stay_physical(ctx):
    """100 steps of the 1,048,576-particle dam break: finite state inside the
    box, bounded cell occupancy, densities near the rest density, a closed
    surface mesh every step."""
    sc = scenes.dam_break_1m()
    p = sc.params()
    m = sc.mc()
    ctx.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
    for _ in range(10):
        ctx.step(p, 10)
        ctx.mc_sample(p, m)
        ctx.mc_classify_count(m)
        ctx.mc_emit(p, m)
        ntri = ctx.mc_count()
        assert 10_000 < ntri < 16_000_000
    g = ctx.download()
    assert np.isfinite(g["position"]).all() and np.isfinite(g["velocity"]).all()
    assert (g["position"] >= 0).all() and (g["position"] <= np.array([4, 2, 2], np.float32)).all()
    assert np.abs(g["velocity"]).max() < 50.0
    table = ctx.read(LF_ARR_TABLE)
    assert np.diff(table.astype(np.int64)).max() < 64
    ctx.lambda_pass(p)
    rho = ctx.read(LF_ARR_DENSITY)
    assert 5000 < np.median(rho) < 8500
    assert sorted(g["order"].tolist()) == list(range(sc.n))


def test_cpp_dropin_solver(ctx, tmp_path):
    """hip::SphSolver (include/fluid/hipsph.hpp) driven from C++ like the
    reference application drives ocl::SphSolver: particles bit-exact against
    the oracle's step, the mesh and the query results equal the Python mirror's."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sc = scenes.dam_break_8k()
    rng = np.random.default_rng(9)
    xs = np.zeros(sc.n, dtype=lf.particle_dtype())
    xs["id"] = sc.ids + 7
    xs["type"] = sc.type
    xs["mass"] = sc.mass
    xs["position"] = sc.position
    xs["velocity"] = sc.velocity
    xs["colour"] = rng.integers(0, 2**32, sc.n, dtype=np.uint64).astype(np.uint32)
    fin = tmp_path / "in.bin"
    xs.tofile(fin)
    queries = [(0.3, 0.3, 0.3), (0.7, 0.2, 0.5), (1.9, 0.9, 0.9)]
    args = [os.path.join(root, "libfluid_amd", "bin", "dropin_main"), str(fin),
            str(tmp_path / "p.bin"), str(tmp_path / "t.bin"), str(tmp_path / "q.txt"),
            repr(sc.h), repr(sc.dt), "1.0", "1.0", "100.0", "4", "0", "-9.8", "0",
            "0", "0", "0", "2", "1", "1", "1"] + [repr(v) for q in queries for v in q]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(tmp_path / "p.bin", dtype=lf.particle_dtype())
    tris = np.fromfile(tmp_path / "t.bin", dtype=lf.triangle_dtype())
    qres = [tuple(int(v) for v in line.split()) for line in open(tmp_path / "q.txt")]

    p = sc.params(flags=lf.LF_FLAG_DIFFUSE, xsph_c=0.0)
    o = oracle.Pbf(p)
    o.load(sc.position, sc.velocity, sc.mass, sc.type, xs["colour"])
    o.step()
    order = o.get(LF_ARR_ORDER)
    _eq(out["id"], xs["id"][order], "ids")
    _eq(out["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "positions")
    _eq(out["velocity"].astype(np.float64), o.get(oracle.CALLER_VELOCITY), "velocities")
    _eq(out["colour"], o.get(oracle.CALLER_COLOUR), "colours")

    solver = lf.SphSolver(sc.h)
    cfg = lf.Config(dt=sc.dt, scale=1.0, resolution=1.0, isolevel=100.0, iteration=4,
                    constant_force=sc.gravity, min_bound=sc.min_bound, max_bound=sc.max_bound,
                    queries=[lf.Query(i, q) for i, q in enumerate(queries)])
    pout, res = solver.advance(cfg, xs)
    solver.ctx.close()
    assert len(tris) == len(res.triangles) > 0
    _eq(tris["v"], res.triangles["v"], "C++ vs Python mesh")
    _eq(tris["c"], res.triangles["c"], "C++ vs Python mesh colours")
    assert qres == [(q.id, q.neighbours, q.average_colour) for q in res.queries]
    assert len(qres) >= 2


def test_mesh_arbitrary_field(ctx):
    """lf_mc_field_upload + classify/emit on a sphere field: the GPU mesher of
    an arbitrary scalar field against the oracle's marching cubes."""
    S, r = 57, 0.3
    step = np.float32(1.0) / np.float32(S - 1)
    ax = np.arange(S, dtype=np.float32) * step
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    field = (np.float32(r) - np.sqrt(((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
                                     .astype(np.float32))).astype(np.float32)
    tris = ctx.mesh_field(field, 0.0, (0.0, 0.0, 0.0), (float(step),) * 3, 2.0)
    otris = oracle.mc_emit(field.astype(np.float64), 0.0, (0, 0, 0), (float(step),) * 3, 2.0)
    assert len(tris) == len(otris) > 1000
    _eq(tris["v"], otris["v"], "sphere mesh")


def test_field_written_after_density_sampling(ctx):
    """lf_write of the field right after a density field was sampled on the
    marching-cubes stream: the write waits for the sampling, and
    classification and emission (back on that stream) mesh the written
    field, equal to the oracle's marching cubes of it."""
    sc = scenes.dam_break_8k()
    p = sc.params()
    ctx.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
    ctx.step(p, 1)
    m = sc.mc()
    ctx.mc_sample(p, m)
    S, origin, step = ctx.mc_info()
    axes = [np.float32(origin[k]) + np.arange(S[k], dtype=np.float32) * np.float32(step[k])
            for k in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    d = np.sqrt(((X - 1.0) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2).astype(np.float32))
    field = (np.float32(m.isolevel) * (np.float32(1.5) - d / np.float32(0.25))).astype(np.float32)
    ctx.write(LF_ARR_FIELD, field)
    ctx.mc_classify_count(m)
    ctx.mc_emit(p, m)
    tris = ctx.mc_triangles()
    otris = oracle.mc_emit(field.astype(np.float64), m.isolevel, origin, step, p.scale)
    assert len(tris) == len(otris) > 500
    _eq(tris["v"], otris["v"], "written field mesh")


def _ramp_mesh():
    """Two triangles forming a tilted ramp across the 8k box's floor region,
    plus one far away that never touches (plane beyond 40 units)."""
    a, b, c, d = (0.0, 0.05, 0.0), (2.0, 0.35, 0.0), (2.0, 0.35, 1.0), (0.0, 0.05, 1.0)
    far = ((100.0, 100.0, 100.0), (101.0, 100.0, 100.0), (100.0, 101.0, 100.0))
    return np.array([(a, b, c), (a, c, d), far], np.float32)


@pytest.mark.parametrize("steps", [1, 5])
def test_mesh_collider_bit_exact(ctx, steps):
    """collideTriangle2 inside every delta pass (oclsph_collision.h:25-58,
    oclsph_kernel.h:364-368): positions and the reflected velocities."""
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    o = _pair(ctx, sc, p)
    mesh = _ramp_mesh()
    ctx.set_colliders(mesh)
    o.set_colliders(mesh)
    try:
        ctx.step(p, steps)
        o.step(steps)
        g = ctx.download()
        _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "position")
        _eq(g["velocity"].astype(np.float64), o.get(oracle.CALLER_VELOCITY), "velocity")
    finally:
        ctx.set_colliders(np.zeros((0, 3, 3), np.float32))
    # the ramp moved particles: the same steps without it differ
    o2 = _pair(ctx, sc, p)
    ctx.step(p, steps)
    assert not np.array_equal(ctx.download()["position"], g["position"])


@pytest.mark.parametrize("steps", [1, 4])
def test_mesh_collider_grid_bit_exact(ctx, steps):
    """A 1,280-triangle sphere cutting into the block plus the ramp: the delta
    pass looks triangles up through the collider grid (per-cell lists of the
    triangles whose +-40 prism may reach the cell); the oracle scans the whole
    list. First hits, positions and velocities must agree exactly."""
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    o = _pair(ctx, sc, p)
    mesh = np.concatenate([_ramp_mesh(), scenes.icosphere((0.9, 0.35, 0.5), 0.25, 3)])
    assert len(mesh) == 3 + 1280
    ctx.set_colliders(mesh)
    o.set_colliders(mesh)
    try:
        ctx.profile(True)
        ctx.profile_reset()
        ctx.step(p, steps)
        stats = ctx.kernel_stats()
        ctx.profile(False)
        assert "collider_grid_fill" in stats  # the grid was built and is in use
        o.step(steps)
        g = ctx.download()
        _eq(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "position")
        _eq(g["velocity"].astype(np.float64), o.get(oracle.CALLER_VELOCITY), "velocity")
    finally:
        ctx.set_colliders(np.zeros((0, 3, 3), np.float32))
