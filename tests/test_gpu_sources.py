"""GPU: sources, drains and spatial queries applied on the device inside
lf_pbf_step (oclsph.hpp:675-706 and 770-791) against the same step fed by
the host-side pre-pass (sources and drains applied to downloaded arrays,
uploaded again), bit-exact; query answers against a numpy restatement over
the host path's sorted state."""
import types

import numpy as np
import pytest

import libfluid_amd as lf
from libfluid_amd import scenes
from libfluid_amd._abi import LF_ARR_ORDER, LF_ARR_TABLE, LF_FLUID, particle_dtype

pytestmark = pytest.mark.gpu

SOURCES = [lf.Source(centre=(1.5, 0.8, 0.5), velocity=(-0.5, 0.0, 0.0), rate=18, tag=7,
                     colour=0xFF00FF00),
           lf.Source(centre=(1.2, 0.6, 0.3), velocity=(0.0, -0.25, 0.1), rate=9, tag=8,
                     colour=0x80FF0040)]
DRAINS = [lf.Drain(centre=(0.25, 0.1, 0.25), width=0.15, depth=0.1),
          lf.Drain(centre=(1.5, 0.1, 0.5), width=0.12, depth=0.1)]
QUERIES = np.array([[0.5, 0.5, 0.5], [0.2, 0.2, 0.2], [1.9, 0.9, 0.9], [0.05, 0.05, 0.9]],
                   np.float32)


def _structured(d):
    xs = np.zeros(len(d["mass"]), dtype=particle_dtype())
    xs["type"] = d["type"]
    xs["mass"] = d["mass"]
    xs["colour"] = d["colour"]
    xs["position"] = d["position"]
    xs["velocity"] = d["velocity"]
    return xs


def _expected_queries(ctx, host_xs, h, scale):
    """Own-cell fluid count and mean ARGB (truncated) over the pre-diffusion
    colours in the step's sorted order."""
    G, _, mn = ctx.grid_info()
    table = ctx.read(LF_ARR_TABLE)
    order = ctx.read(LF_ARR_ORDER)
    col = host_xs["colour"][order]
    typ = host_xs["type"][order]
    counts, colours = [], []
    for q in QUERIES:
        cells = []
        for k in range(3):
            r = (np.float32(q[k]) / np.float32(scale) - np.float32(mn[k])) / np.float32(h)
            cells.append(int(r) & 1023 if r >= 0 else 0)
        z = lf._morton(*cells)
        n, packed = 0, 0
        if z + 1 < G:
            rng = slice(int(table[z]), int(table[z + 1]))
            c = col[rng][typ[rng] == LF_FLUID].astype(np.uint64)
            n = len(c)
            if n:
                argb = np.stack([(c >> 24) & 255, (c >> 16) & 255, (c >> 8) & 255, c & 255],
                                axis=1).astype(np.float32)
                avg = argb.sum(axis=0) / np.float32(n)
                for a in avg:
                    packed = (packed << 8) | (int(a) & 255)
        counts.append(n)
        colours.append(packed)
    return np.array(counts, np.uint32), np.array(colours, np.uint32)


def test_sources_drains_queries_match_host_path():
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE, xsph_c=0.01)
    cfg = types.SimpleNamespace(sources=SOURCES, drains=DRAINS, scale=sc.scale)
    dev = lf.Context(0)
    host = lf.Context(0)
    try:
        dev.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        host.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        dev.set_sources(SOURCES)
        dev.set_drains(DRAINS)
        dev.set_queries(QUERIES)
        sizes = []
        for step in range(6):
            xs = lf._apply_sources_drains(cfg, _structured(host.download()), sc.h)
            host.upload(xs["position"], xs["velocity"], xs["mass"], xs["type"], xs["colour"])
            host.step(p)
            dev.step(p)
            assert dev.n == host.n == len(xs), "step %d: sizes %d %d %d" % (step, dev.n, host.n,
                                                                          len(xs))
            sizes.append(dev.n)
            a, b = dev.download(), host.download()
            for k in ("position", "velocity", "mass", "type", "colour"):
                bad = a[k] != b[k]
                assert not bad.any(), "step %d %s: %d differ" % (step, k, int(bad.sum()))
            counts, colours = dev.query_results()
            ec, ecol = _expected_queries(host, xs, sc.h, sc.scale)
            assert (counts == ec).all(), "step %d counts %s vs %s" % (step, counts, ec)
            assert (colours == ecol).all(), "step %d colours %s vs %s" % (step, colours, ecol)
        # the sources add 4x5 + 3x3 = 29 particles per step and the drains remove some
        assert sizes[0] != len(sc.mass) and len(set(sizes)) > 1
        assert max(int(c) for c in counts) > 0
        # a batch upload is refused while sources are set; clearing them allows it
        with pytest.raises(lf.LfError):
            dev.batch_upload([p], [dict(position=sc.position, velocity=sc.velocity,
                                        mass=sc.mass)])
        dev.set_sources([])
        dev.set_drains([])
        dev.set_queries(np.zeros((0, 3), np.float32))
        dev.step(p)
    finally:
        dev.close()
        host.close()


def test_query_results_need_a_step():
    sc = scenes.small_block(n_side=6)
    p = sc.params()
    with lf.Context(0) as c:
        c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        c.set_queries(QUERIES[:2])
        with pytest.raises(lf.LfError):
            c.query_results()
        c.step(p)
        counts, colours = c.query_results()
        assert counts.shape == (2,)


def test_drains_empty_the_scene_and_sources_refill():
    """A drain that takes every particle leaves an empty state whose step and
    queries still work (the reference's depleted case, oclsph.hpp:711-715);
    a source then refills it."""
    sc = scenes.small_block(n_side=6)
    p = sc.params()
    with lf.Context(0) as c:
        c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        c.set_drains([lf.Drain(centre=(0.5, 0.5, 0.5), width=10.0, depth=0.0)])
        c.set_queries(QUERIES[:2])
        c.step(p)
        assert c.n == 0
        counts, colours = c.query_results()
        assert (counts == 0).all() and (colours == 0).all()
        c.step(p)  # an empty step
        assert c.n == 0
        c.set_drains([])
        c.set_sources([lf.Source(centre=(0.5, 0.5, 0.5), velocity=(0.0, 0.0, 0.0), rate=16,
                                 tag=1, colour=0xFFFFFFFF)])
        c.step(p)
        assert c.n == 16
        d = c.download()
        assert np.isfinite(d["position"]).all() and (d["colour"] == 0xFFFFFFFF).all()
        counts, _ = c.query_results()
        assert counts.shape == (2,)
