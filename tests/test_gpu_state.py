"""GPU: state handling across calls — buffer changes under a captured graph,
what the autotuner may and may not touch, and the ordering of the density
field's read of p* on the marching-cubes stream against later in-place writes
of p* on the main stream. Each case is checked against the same calls made
without the hazard (bit for bit)."""
import numpy as np
import pytest

import libfluid_amd as lf
from libfluid_amd import scenes
import oracle
from libfluid_amd._abi import LF_ARR_FIELD, LF_ARR_ORDER, LF_ERR_DIVERGED, LF_ERR_STATE

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = a != b
    if bad.any():
        i = np.argwhere(bad)[0]
        raise AssertionError("%s: %d differ, first %s" % (what, int(bad.sum()), i.tolist()))


def test_capacity_change_drops_captured_graph():
    """A graph captured with marching cubes holds the triangle buffers'
    addresses: changing the capacity frees them, so the graph must go with
    them; eager meshing afterwards is unaffected."""
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    m = sc.mc()
    with lf.Context(0) as g, lf.Context(0) as e:
        for c in (g, e):
            c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
            c.step(p, 1)
            c.mesh(p, m)
        g.graph_capture(2, m)
        g.graph_launch(1)
        e.step(p, 1)
        e.mesh(p, m)
        e.step(p, 1)
        e.mesh(p, m)
        g.mc_set_capacity(1 << 20)
        with pytest.raises(lf.LfError) as err:
            g.graph_launch(1)
        assert err.value.code == LF_ERR_STATE
        with pytest.raises(lf.LfError):  # the triangles went with the buffers
            g.mc_triangles()
        for c in (g, e):
            c.step(p, 1)
        tg, te = g.mesh(p, m), e.mesh(p, m)
        assert tg.shape == te.shape and len(tg) > 1000
        _same(tg["v"], te["v"], "triangles after the capacity change")
        _same(g.download()["position"], e.download()["position"], "positions")


def test_autotune_refuses_sources_and_keeps_query_answers():
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    with lf.Context(0) as c:
        c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        c.step(p, 1)
        c.set_sources([lf.Source(centre=(1.5, 0.8, 0.5), velocity=(0, 0, 0), rate=9, tag=1,
                                 colour=0xFFFFFFFF)])
        with pytest.raises(lf.LfError) as err:
            c.autotune(1)
        assert err.value.code == LF_ERR_STATE
        c.set_sources([])
        c.set_drains([lf.Drain(centre=(0.2, 0.1, 0.2), width=0.1, depth=0.0)])
        with pytest.raises(lf.LfError):
            c.autotune(1)
        c.set_drains([])
        c.set_queries(np.array([[0.3, 0.2, 0.3], [0.6, 0.1, 0.4]], np.float32))
        c.step(p, 1)
        counts, colours = c.query_results()
        assert counts.max() > 0
        n_before = c.download()["position"].shape[0]
        c.autotune(1)
        counts2, colours2 = c.query_results()
        _same(counts2, counts, "query counts after autotune")
        _same(colours2, colours, "query colours after autotune")
        assert c.download()["position"].shape[0] == n_before


def test_density_field_sampled_mid_step_then_delta():
    """Stage by stage: the density field of the p* after the first delta pass,
    sampled on the marching-cubes stream, then the next lambda and delta passes
    (delta rewrites p* in place on the main stream). The field must be the one
    of the p* it was sampled from, as when the host waits after sampling."""
    sc = scenes.dam_break_1m()
    p = sc.params()
    m = sc.mc()
    fields = []
    for wait in (True, False):
        with lf.Context(0) as c:
            c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
            c.step(p, 3)
            c.grid_build(p)
            c.lambda_pass(p)
            c.delta_pass(p)
            c.mc_sample(p, m)
            if wait:
                c.sync()
            c.lambda_pass(p)
            c.delta_pass(p)
            c.finalise(p)
            fields.append(c.read(LF_ARR_FIELD))
    assert (fields[0] > 0).sum() > 10000
    _same(fields[1], fields[0], "density field sampled before an in-place delta pass")


def test_guard_stops_a_non_finite_state():
    """Five particles with NaN positions: the key pass counts them, the pair
    passes of that step do nothing (the state stays the step's start, sorted),
    and the step calls report LF_ERR_DIVERGED from then on; a new upload
    clears it."""
    sc = scenes.dam_break_8k()
    p = sc.params()
    pos = sc.position.copy()
    pos[[10, 500, 4000, 7000, 7999]] = np.nan
    with lf.Context(0) as c:
        c.upload(pos, sc.velocity, sc.mass, sc.type, sc.colour)
        c.step(p, 1)
        with pytest.raises(lf.LfError) as err:
            c.sync()
        assert err.value.code == LF_ERR_DIVERGED and "5 particle" in str(err.value)
        assert c.health() == (5, 0, 0)
        with pytest.raises(lf.LfError) as err:
            c.step(p, 1)
        assert err.value.code == LF_ERR_DIVERGED
        g = c.download()
        order = g["order"].astype(np.int64)
        finite = np.isfinite(pos[order]).all(1)
        assert finite.sum() == sc.n - 5
        _same(g["position"][finite], pos[order][finite], "positions of the halted step")
        c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        c.step(p, 2)
        c.sync()
        assert c.health() == (0, 0, 0)


def test_guard_overfull_cell_and_large_bucket_order():
    """400 coincident particles in one cell: above the default guard (256) the
    step reports the cell; with the guard off the step runs and equals the
    oracle bit for bit (the cell's bucket is ordered by the large-bucket sort)."""
    sc = scenes.small_block(n_side=8)
    pos = np.concatenate([sc.position, np.tile(sc.position[100:101], (400, 1))])
    n = len(pos)
    vel = np.zeros_like(pos)
    mass = np.ones(n, np.float32)
    typ = np.zeros(n, np.int32)
    col = np.full(n, 0xFF102030, np.uint32)
    p = sc.params()
    with lf.Context(0) as c:
        c.upload(pos, vel, mass, typ, col)
        c.step(p, 1)
        with pytest.raises(lf.LfError) as err:
            c.sync()
        assert err.value.code == LF_ERR_DIVERGED
        h = c.health()
        assert h[0] == 0 and h[1] == 1 and h[2] >= 401
        c.set_guard(0)
        c.upload(pos, vel, mass, typ, col)
        c.step(p, 1)
        c.sync()
        g = c.download()
    o = oracle.Pbf(p)
    o.load(pos, vel, mass, typ, col)
    o.step()
    _same(g["order"], o.get(LF_ARR_ORDER), "sorted order")  # one step: upload index = pre-step index
    _same(g["position"].astype(np.float64), o.get(oracle.CALLER_POSITION), "positions")
