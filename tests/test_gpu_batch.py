"""GPU: batched scenes, execution configurations, HIP-graph replay and the
autotuner, each checked bit for bit against the plain single-scene path
(itself bit-exact against the oracle in test_gpu_parity.py).

A batch steps several scenes in the same launches; each scene's particles
must come out exactly as when that scene is stepped alone, so states are
compared per particle (by upload index). Execution configurations (workgroup
size, persistent work queue, neighbour-list skin), graph replay and the
autotuner change speed only, never results.
"""
import numpy as np
import pytest

import libfluid_amd as lf
import oracle
from libfluid_amd import scenes
from libfluid_amd._abi import LF_ARR_CELLKEY, LF_ARR_FIELD, LF_ARR_KEYS, LF_ARR_TABLE

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = lf.Context(0)
    yield c
    c.close()


def _by_upload_index(state, offset=0):
    """State arrays re-indexed by upload index (minus offset)."""
    order = state["order"].astype(np.int64) - offset
    out = {}
    for k in ("position", "velocity", "colour", "type"):
        a = np.empty_like(state[k])
        a[order] = state[k]
        out[k] = a
    return out


def _same(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    bad = a != b
    if a.dtype.kind == "f":
        bad &= ~(np.isnan(a) & np.isnan(b))
    if bad.any():
        i = np.argwhere(bad)[0]
        raise AssertionError("%s: %d differ, first %s: %s vs %s" % (
            what, int(bad.sum()), i.tolist(), a[tuple(i)], b[tuple(i)]))


def _small_batch(n_scenes=6, seed=11):
    sc = scenes.batch_scenes(n_scenes, lo=500, hi=12000, seed=seed)
    for i, s in enumerate(sc):  # per-scene constants differ: XSPH on every other scene
        s.xsph_c = 0.01 if i % 2 == 0 else 0.0
    return sc


def _single(sc, steps, flags=lf.LF_FLAG_DIFFUSE):
    with lf.Context(0) as c:
        c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        c.step(sc.params(flags=flags), steps)
        return _by_upload_index(c.download())


def _batch_states(c, sc):
    g = c.download()
    off = c.batch_info()
    out = []
    for s in range(len(sc)):
        part = {k: v[off[s]:off[s + 1]] for k, v in g.items()}
        out.append(_by_upload_index(part, off[s]))
    return out


def _upload_batch(c, sc, flags=lf.LF_FLAG_DIFFUSE):
    c.batch_upload([s.params(flags=flags) for s in sc],
                   [dict(position=s.position, velocity=s.velocity, mass=s.mass, type=s.type,
                         colour=s.colour) for s in sc])


def test_batch_equals_scenes_alone(ctx):
    sc = _small_batch()
    steps = 3
    _upload_batch(ctx, sc)
    ctx.batch_step(steps)
    got = _batch_states(ctx, sc)
    for s, scene in enumerate(sc):
        ref = _single(scene, steps)
        for k in ("position", "velocity", "colour"):
            _same(got[s][k], ref[k], "scene %d %s" % (s, k))


def test_batch_keys_and_tables(ctx):
    sc = _small_batch(4, seed=12)
    _upload_batch(ctx, sc)
    ctx.batch_step(1)
    keys = ctx.read(LF_ARR_KEYS)
    assert np.all(np.diff(keys.astype(np.int64)) >= 0), "sort keys not sorted"
    off = ctx.batch_info()
    cell = ctx.read(LF_ARR_CELLKEY)
    base = 0
    for s, scene in enumerate(sc):
        ctx.batch_select(s)
        G, _, _ = ctx.grid_info()
        k = keys[off[s]:off[s + 1]].astype(np.int64)
        assert k.min() >= base and k.max() <= base + G, "scene %d keys out of its range" % s
        # the sort key is the clamped local key plus the scene base
        _same(k - base, np.minimum(cell[off[s]:off[s + 1]], G).astype(np.int64),
              "scene %d local keys" % s)
        base += G + 1
    table = ctx.read(LF_ARR_TABLE, count=base)
    assert table[0] == 0 and np.all(np.diff(table.astype(np.int64)) >= 0)


@pytest.mark.parametrize("threads,persistent,skin,ppt", [
    (128, False, 0.2, 1), (512, False, 0.2, 1), (256, True, 0.2, 1), (128, True, 0.1, 1),
    (256, False, 0.0, 1), (512, True, 0.3, 1), (128, False, 0.2, 2), (256, False, 0.3, 4),
    (512, False, 0.1, 4)])
def test_exec_configs_same_results(ctx, threads, persistent, skin, ppt):
    sc = _small_batch(5, seed=13)
    _upload_batch(ctx, sc)
    ctx.batch_step(2)
    ref = _batch_states(ctx, sc)
    with lf.Context(0) as c:
        c.exec_config(threads, persistent, skin, particles_per_thread=ppt)
        _upload_batch(c, sc)
        c.batch_step(2)
        got = _batch_states(c, sc)
    for s in range(len(sc)):
        for k in ("position", "velocity", "colour"):
            _same(got[s][k], ref[s][k], "cfg %s scene %d %s" % ((threads, persistent, skin), s, k))


def test_single_scene_exec_configs_bit_exact():
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    ref = _single(sc, 4)
    for cfg in [(128, True, 0.1), (512, False, 0.0)]:
        with lf.Context(0) as c:
            c.exec_config(*cfg)
            c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
            c.step(p, 4)
            got = _by_upload_index(c.download())
        for k in ("position", "velocity", "colour"):
            _same(got[k], ref[k], "cfg %s %s" % (cfg, k))


def test_graph_replays_match_eager():
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    m = sc.mc()
    with lf.Context(0) as g, lf.Context(0) as e:
        for c in (g, e):
            c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
            c.step(p, 1)
            c.mesh(p, m)  # allocates the marching-cubes buffers before capture
        g.graph_capture(2, m)
        for replay in range(3):
            g.graph_launch(1)
            e.step(p, 1)
            e.mesh(p, m)
            e.step(p, 1)
            tri_e = e.mesh(p, m)
            sg, se = g.download(), e.download()
            for k in ("position", "velocity", "colour", "order"):
                _same(sg[k], se[k], "replay %d %s" % (replay, k))
            _same(g.read(LF_ARR_FIELD), e.read(LF_ARR_FIELD), "replay %d field" % replay)
            tri_g = g.mc_triangles()
            assert tri_g.shape == tri_e.shape
            _same(tri_g["v"], tri_e["v"], "replay %d triangles" % replay)


def test_graph_persistent_batch_matches_eager():
    sc = _small_batch(5, seed=14)
    with lf.Context(0) as g, lf.Context(0) as e:
        for c in (g, e):
            c.exec_config(256, True, 0.2)
            _upload_batch(c, sc)
            c.batch_step(1)
        g.graph_capture(4)
        g.graph_launch(2)
        e.batch_step(8)
        for s, (a, b) in enumerate(zip(_batch_states(g, sc), _batch_states(e, sc))):
            for k in ("position", "velocity", "colour"):
                _same(a[k], b[k], "scene %d %s" % (s, k))


def test_graph_needs_even_steps_and_a_step(ctx):
    sc = scenes.small_block(n_side=8)
    with lf.Context(0) as c:
        c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        with pytest.raises(lf.LfError):
            c.graph_capture(2)  # no eager step yet
        c.step(sc.params(), 1)
        with pytest.raises(lf.LfError):
            c.graph_capture(3)
        with pytest.raises(lf.LfError):
            c.graph_launch(1)  # nothing captured


def test_autotune_keeps_state():
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    ref = _single(sc, 3)
    with lf.Context(0) as c:
        c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        c.step(p, 1)
        before = c.download()
        r = c.autotune(2)
        assert r["trials"] == 48 and r["threads"] in (128, 256, 512) and r["ms_per_step"] > 0
        assert r["particles_per_thread"] in (1, 2, 4) and r["mc_tile"] == 0 and r["mc_ms"] == 0
        after = c.download()
        for k in ("position", "velocity", "colour", "order"):
            _same(after[k], before[k], "autotune changed %s" % k)
        c.step(p, 2)
        got = _by_upload_index(c.download())
    for k in ("position", "velocity", "colour"):
        _same(got[k], ref[k], "after autotune %s" % k)


def test_batch_density_mesh_per_scene(ctx):
    sc = _small_batch(3, seed=15)
    _upload_batch(ctx, sc)
    ctx.batch_step(2)
    for s, scene in enumerate(sc):
        m = scene.mc()
        ctx.batch_select(s)
        ctx.mc_sample(None, m)
        ctx.mc_classify_count(m)
        ctx.mc_emit(None, m)
        field = ctx.read(LF_ARR_FIELD)
        tris = ctx.mc_triangles()
        with lf.Context(0) as c:
            p = scene.params(flags=lf.LF_FLAG_DIFFUSE)
            c.upload(scene.position, scene.velocity, scene.mass, scene.type, scene.colour)
            c.step(p, 2)
            ref_tris = c.mesh(p, m)
            ref_field = c.read(LF_ARR_FIELD)
        _same(field, ref_field, "scene %d density field" % s)
        assert tris.shape == ref_tris.shape
        _same(tris["v"], ref_tris["v"], "scene %d triangles" % s)


def test_batch_rejects_mixed_iterations(ctx):
    sc = _small_batch(2, seed=16)
    ps = [s.params() for s in sc]
    ps[1].iterations = 3
    with pytest.raises(lf.LfError):
        ctx.batch_upload(ps, [dict(position=s.position, velocity=s.velocity, mass=s.mass)
                              for s in sc])


def test_density_tile_shapes_same_field():
    """The density field's sample tile (a runtime choice: columns of 1, 2 or 4
    samples per thread) changes nothing: the same terms in the same order per
    sample, whether staged in LDS or read from global memory, on a cubic
    lattice and on a ragged one whose tiles are cut at every edge (equal to
    the oracle's FP32 field there too). The classification then skips the
    blocks whose samples all lie in tiles written as zero: the triangles equal
    the oracle's marching cubes of the field for every tile shape. The
    autotuner times the shapes on the last density lattice."""
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    lattices = (sc.mc(), lf.mc_density_params(sc.isolevel, (61, 45, 37)))
    fields, tris = {}, {}
    for tile in range(4):
        with lf.Context(0) as c:
            c.exec_config(256, False, 0.2, mc_tile=tile)
            c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
            c.step(p, 2)
            for j, m in enumerate(lattices):
                c.mc_sample(p, m)
                c.mc_classify_count(m)  # skips blocks in tiles the field wrote as zero
                c.mc_emit(p, m)
                fields[tile, j] = c.read(LF_ARR_FIELD)
                tris[tile, j] = c.mc_triangles()
            if tile == 2:
                r = c.autotune(1)
                assert r["trials"] == 52 and r["mc_tile"] in range(4) and r["mc_ms"] > 0
    o = oracle.Pbf(p)
    o.load(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
    o.step(2)
    ragged = o.mc_density(lattices[1].samples)
    for j in range(2):
        assert (fields[0, j] > 0).sum() > 1000
        for tile in (1, 2, 3):
            _same(fields[tile, j], fields[0, j], "tile %d, lattice %d" % (tile, j))
    _same(fields[0, 1].astype(np.float64).ravel(), ragged.ravel(), "ragged lattice against the oracle")
    for j, m in enumerate(lattices):
        origin, step = o.density_lattice(m.samples)
        want = oracle.mc_emit(fields[0, j].astype(np.float64).reshape(tuple(m.samples)), m.isolevel,
                              origin, step, 1.0)
        assert len(want) > 100
        for tile in range(4):
            _same(tris[tile, j]["v"], want["v"], "triangles, tile %d, lattice %d" % (tile, j))
