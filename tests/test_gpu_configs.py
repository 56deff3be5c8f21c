"""GPU: the benchmark configurations at their full sizes.

* 64 batched scenes, log-uniform 2k-200k particles (seed 3), FP32, autotuned,
  whole steps replayed from a HIP graph: the smallest, the largest and two more
  scenes bit for bit against each scene stepped alone; every scene's particle
  count, mass and finiteness, and its particles inside its own box.
* 32 scenes of ~8M particles in total (log-uniform 50k-500k, seed 5),
  FAST_MATH + KAHAN, 6 iterations, graph-replayed: the largest and smallest
  scenes against the scene stepped alone (bit for bit: a batch is the same
  arithmetic as its scenes), and the largest scene's 512^3 density field and
  mesh against the single-scene path.
* The 4,194,304-particle double dam break, 20 steps: the error against the
  FP64 oracle one step at a time from the oracle's own state (teacher forced)
  and accumulated over the 20 steps: plain FP32, mixed precision (FAST_MATH +
  KAHAN) and mixed with the 16-bit cell-relative neighbour positions.
"""
import numpy as np
import pytest

import libfluid_amd as lf
import oracle
from libfluid_amd import scenes
from libfluid_amd._abi import LF_ARR_DENSITY, LF_ARR_FIELD, LF_ARR_LAMBDA, LF_ARR_ORDER

pytestmark = pytest.mark.gpu

MIXED = lf.LF_FLAG_FAST_MATH | lf.LF_FLAG_KAHAN
PACKED = MIXED | lf.LF_FLAG_PACKED16


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = a != b
    if a.dtype.kind == "f":
        bad &= ~(np.isnan(a) & np.isnan(b))
    if bad.any():
        i = np.argwhere(bad)[0]
        raise AssertionError("%s: %d differ, first %s" % (what, int(bad.sum()), i.tolist()))


def _by_upload_index(state, offset=0):
    order = state["order"].astype(np.int64) - offset
    out = {}
    for k in ("position", "velocity", "colour"):
        a = np.empty_like(state[k])
        a[order] = state[k]
        out[k] = a
    return out


def _batch(sc, flags):
    c = lf.Context(0)
    c.batch_upload([s.params(flags=flags) for s in sc],
                   [dict(position=s.position, velocity=s.velocity, mass=s.mass, type=s.type,
                         colour=s.colour) for s in sc])
    return c


def _alone(scene, flags, steps):
    with lf.Context(0) as c:
        c.upload(scene.position, scene.velocity, scene.mass, scene.type, scene.colour)
        c.step(scene.params(flags=flags), steps)
        return _by_upload_index(c.download())


def _step_batch_like_the_bench(c):
    """One eager step, the autotuner, one eager step, a 2-step graph replayed
    twice: 6 steps."""
    c.batch_step(1)
    tune = c.autotune(1)
    c.batch_step(1)
    c.graph_capture(2)
    c.graph_launch(2)
    c.sync()
    print("autotune", tune, flush=True)
    return 6


def test_c3_64_scenes_graph_autotuned():
    sc = scenes.batch_scenes(64)
    counts = np.array([s.n for s in sc])
    assert len(sc) == 64 and counts.min() >= 1500 and counts.max() <= 260000
    c = _batch(sc, 0)
    try:
        steps = _step_batch_like_the_bench(c)
        g = c.download()
        off = c.batch_info()
        diag = [c.diagnostics(s) for s in range(len(sc))]
    finally:
        c.close()
    for s, scene in enumerate(sc):
        part = {k: v[off[s]:off[s + 1]] for k, v in g.items()}
        assert len(part["position"]) == scene.n
        pos = part["position"]
        assert np.isfinite(pos).all() and np.isfinite(part["velocity"]).all(), s
        lo, hi = np.asarray(scene.min_bound, np.float32), np.asarray(scene.max_bound, np.float32)
        assert (pos >= lo).all() and (pos <= hi).all(), "scene %d left its box" % s
        assert diag[s]["fluid"] == scene.n and diag[s]["mass"] == float(scene.n)
    pick = sorted({int(np.argmin(counts)), int(np.argmax(counts)), 17, 42})
    for s in pick:
        part = {k: v[off[s]:off[s + 1]] for k, v in g.items()}
        got = _by_upload_index(part, off[s])
        ref = _alone(sc[s], 0, steps)
        for k in ("position", "velocity", "colour"):
            _same(got[k], ref[k], "scene %d (%d particles) %s" % (s, sc[s].n, k))


def test_c5_8m_mixed_6_iterations_with_512_mesh():
    sc = scenes.combined_scenes()
    counts = np.array([s.n for s in sc])
    assert len(sc) == 32 and abs(counts.sum() - 8_000_000) < 80_000
    assert all(s.iterations == 6 for s in sc)
    big, small = int(np.argmax(counts)), int(np.argmin(counts))
    m = sc[big].mc(attributes=0)
    assert tuple(m.samples) == (512, 512, 512)
    c = _batch(sc, MIXED)
    try:
        c.mc_set_capacity(48 << 20)
        steps = _step_batch_like_the_bench(c)
        g = c.download()
        off = c.batch_info()
        c.batch_select(big)
        c.mc_sample(None, m)
        c.mc_classify_count(m)
        c.mc_emit(None, m)
        field = c.read(LF_ARR_FIELD)
        tris = c.mc_triangles()
    finally:
        c.close()
    assert np.isfinite(g["position"]).all()
    for s in (big, small):
        part = {k: v[off[s]:off[s + 1]] for k, v in g.items()}
        got = _by_upload_index(part, off[s])
        ref = _alone(sc[s], MIXED, steps)
        for k in ("position", "velocity", "colour"):
            _same(got[k], ref[k], "scene %d (%d particles) %s" % (s, sc[s].n, k))
    scene = sc[big]
    p = scene.params(flags=MIXED)
    with lf.Context(0) as a:
        a.mc_set_capacity(48 << 20)
        a.upload(scene.position, scene.velocity, scene.mass, scene.type, scene.colour)
        a.step(p, steps)
        ref_tris = a.mesh(p, m)
        ref_field = a.read(LF_ARR_FIELD)
    assert field.size == 512 ** 3 and (field > 0).sum() > 100_000
    _same(field, ref_field, "512^3 density field of scene %d" % big)
    assert len(tris) > 100_000 and tris.shape == ref_tris.shape
    _same(tris["v"], ref_tris["v"], "triangles of scene %d" % big)


def _by_id(values, ids):
    out = np.empty_like(values)
    out[ids.astype(np.int64)] = values
    return out


def _err(got, ref, fluid, sc):
    """Per-particle errors: density relative, lambda relative to max |lambda|,
    position in units of h."""
    rho, lam, pos = got
    rrho, rlam, rpos = ref
    return (np.abs(rho[fluid] - rrho[fluid]) / rrho[fluid],
            np.abs(lam[fluid] - rlam[fluid]) / np.max(np.abs(rlam[fluid])),
            np.max(np.abs(pos - rpos), axis=1) / sc.scale / sc.h)


def _summary(e):
    """(max, 99.99th percentile, median) of each error."""
    return [(float(x.max()), float(np.quantile(x, 0.9999)), float(np.median(x))) for x in e]


def test_c4_4m_errors_over_20_steps():
    """Teacher forced: each step starts the GPU from the FP64 oracle's state
    (rounded to FP32) and is compared with the oracle's next state. Accumulated:
    the GPU runs 20 steps on its own from the common start. Density relative,
    lambda relative to max |lambda|, position in units of h; the maximum and the
    99.99th percentile over the particles.

    Bars: in every teacher-forced step the 99.99th percentile (all but ~400 of
    the 4M particles) of the density error <= 1e-4 (the mixed-precision target)
    for every mode, and each mode's largest error within twice plain FP32's
    plus 1e-4. The maximum itself is reported, not bounded: in one of the 20
    steps (the 7th) a cluster of ~10 particles has density errors of 6e-4 to
    5e-3 in every FP32 evaluation, the CPU restatement in FP32 included (the
    plain FP32 path is bit-identical to it), so it is the conditioning of the
    reference's step at that state, not the precision options. Accumulated over
    20 steps the FP32 and FP64 runs of this collapse drift apart; the medians
    are reported and the options must stay within 3x plain FP32's."""
    sc = scenes.double_dam_break_4m()
    steps = 20
    fluid = sc.type == lf.LF_FLUID
    o = oracle.Pbf(sc.params(flags=lf.LF_FLAG_DIFFUSE), fp64=True)
    o.load(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
    ids = np.arange(sc.n, dtype=np.int64)  # oracle caller order -> upload index
    modes = ((0, "plain FP32"), (MIXED, "mixed"), (PACKED, "mixed + packed16"))
    steps_err = {f: [] for f, _ in modes}
    ctxs = {f: lf.Context(0) for f, _ in modes}

    def state(c):
        st = c.download()
        up = st["order"]
        return (_by_id(c.read(LF_ARR_DENSITY).astype(np.float64), up),
                _by_id(c.read(LF_ARR_LAMBDA)[:, 3].astype(np.float64), up),
                _by_id(st["position"].astype(np.float64), up))

    try:
        for k in range(steps):
            x0 = o.get(oracle.CALLER_POSITION).astype(np.float32)
            v0 = o.get(oracle.CALLER_VELOCITY).astype(np.float32)
            m0 = o.get(oracle.CALLER_MASS).astype(np.float32)
            col0 = o.get(oracle.CALLER_COLOUR)
            t0 = sc.type[ids]
            o.step(1)
            order = o.get(LF_ARR_ORDER).astype(np.int64)  # pre-step index of each sorted particle
            ref = (_by_id(o.get(LF_ARR_DENSITY), order), _by_id(o.get(LF_ARR_LAMBDA), order),
                   _by_id(o.get(oracle.CALLER_POSITION), order))
            for flags, c in ctxs.items():
                c.upload(x0, v0, m0, t0, col0)
                c.step(sc.params(flags=flags | lf.LF_FLAG_DIFFUSE), 1)
                steps_err[flags].append(_summary(_err(state(c), ref, t0 == lf.LF_FLUID, sc)))
            ids = ids[order]
            print("step %2d teacher-forced density (max, 99.99%%): %s" % (
                k + 1, "; ".join("%s %.2e %.2e" % ((tag,) + steps_err[f][-1][0][:2])
                                 for f, tag in modes)), flush=True)
        ref = (_by_id(o.get(LF_ARR_DENSITY), ids), _by_id(o.get(LF_ARR_LAMBDA), ids),
               _by_id(o.get(oracle.CALLER_POSITION), ids))
        acc = {}
        for flags, c in ctxs.items():
            c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
            c.step(sc.params(flags=flags | lf.LF_FLAG_DIFFUSE), steps)
            acc[flags] = _summary(_err(state(c), ref, fluid, sc))
    finally:
        for c in ctxs.values():
            c.close()
    for flags, tag in modes:
        w = [tuple(max(st[q][i] for st in steps_err[flags]) for i in range(3)) for q in range(3)]
        print("4M, 20 steps, %s: teacher-forced worst step (max / 99.99%% / median) density "
              "%.2e / %.2e / %.2e, lambda %.2e / %.2e / %.2e, position %.2e / %.2e / %.2e of h; "
              "accumulated density %.2e / %.2e / %.2e, lambda %.2e / %.2e / %.2e, position "
              "%.2e / %.2e / %.2e of h" % ((tag,) + tuple(v for t in w for v in t)
                                           + tuple(v for t in acc[flags] for v in t)), flush=True)
        for k in range(steps):
            e, e0 = steps_err[flags][k], steps_err[0][k]
            assert e[0][1] <= 1e-4, "step %d %s" % (k + 1, tag)
            assert e[0][0] <= 2 * e0[0][0] + 1e-4, "step %d %s" % (k + 1, tag)
        assert np.isfinite([v for t in acc[flags] for v in t]).all()
        assert acc[flags][0][2] <= 3 * acc[0][0][2] + 1e-6
        assert acc[flags][2][2] <= 3 * acc[0][2][2] + 1e-6
