"""GPU: the mixed-precision options and the FP64 diagnostics.

* LF_FLAG_FAST_MATH (hardware sqrt / reciprocal in the pair terms) with and
  without LF_FLAG_KAHAN, and LF_FLAG_PACKED16 (neighbour positions as 16-bit
  cell-relative fixed point), against the FP64 oracle on the same inputs, one
  step (teacher forced): density relative error <= 1e-4, lambda error <= 1e-4
  of max|lambda| (5e-4 with PACKED16: one quantum, h/65536, per coordinate on
  top of the FP32 rounding of the inputs), position error <= 1e-4 h. The
  plain FP32 path is held to the same bar, and FAST_MATH must not be far worse
  than plain FP32.
* lf_diagnostics: FP64 mass / momentum / density error against numpy FP64
  sums of the downloaded state, and reproducible call to call.
* The density-error convergence check: a tolerance that every scene meets
  after the first lambda pass skips all delta passes (the state equals a
  0-iteration step, bit for bit); an unreachable one changes nothing; in a
  batch, each scene follows its own tolerance.
"""
import numpy as np
import pytest

import libfluid_amd as lf
import oracle
from libfluid_amd import scenes
from libfluid_amd._abi import LF_ARR_DENSITY, LF_ARR_LAMBDA, LF_ARR_ORDER

pytestmark = pytest.mark.gpu

DENSITY_REL_TOL = 1e-4
LAMBDA_TOL = 1e-4      # of max |lambda|
LAMBDA_TOL_PACKED = 5e-4
POSITION_TOL = 1e-4    # of h (solver units)


def _lambda_tol(flags):
    return LAMBDA_TOL_PACKED if flags & lf.LF_FLAG_PACKED16 else LAMBDA_TOL


def _by_id(values, order):
    out = np.empty_like(values)
    out[order.astype(np.int64)] = values
    return out


def _gpu_state(sc, flags, steps=1):
    """Density and lambda of the last lambda pass and positions, by upload index."""
    p = sc.params(flags=flags)
    with lf.Context(0) as c:
        c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        c.step(p, steps)
        st = c.download()
        ids = st["order"]  # upload index of each particle, in the last step's sorted order
        rho = _by_id(c.read(LF_ARR_DENSITY).astype(np.float64), ids)
        lam = _by_id(c.read(LF_ARR_LAMBDA)[:, 3].astype(np.float64), ids)
        pos = _by_id(st["position"].astype(np.float64), ids)
    return rho, lam, pos


def _oracle_state(sc, flags, fp64, steps=1):
    o = oracle.Pbf(sc.params(flags=flags), fp64=fp64)
    o.load(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
    ids = np.arange(sc.n, dtype=np.int64)
    for _ in range(steps):  # compose each step's sort permutation
        o.step(1)
        ids = ids[o.get(LF_ARR_ORDER).astype(np.int64)]
    rho = _by_id(o.get(LF_ARR_DENSITY), ids)
    lam = _by_id(o.get(LF_ARR_LAMBDA), ids)
    pos = _by_id(o.get(oracle.CALLER_POSITION), ids)  # caller arrays: sorted order
    return rho, lam, pos


def _gpu_one_step(sc, flags):
    return _gpu_state(sc, flags, 1)


def _oracle_one_step(sc, flags, fp64):
    return _oracle_state(sc, flags, fp64, 1)


def _errors(got, ref, sc):
    rho, lam, pos = got
    rrho, rlam, rpos = ref
    fluid = sc.type == lf.LF_FLUID
    e_rho = np.max(np.abs(rho[fluid] - rrho[fluid]) / rrho[fluid])
    e_lam = np.max(np.abs(lam[fluid] - rlam[fluid])) / np.max(np.abs(rlam[fluid]))
    e_pos = np.max(np.abs(pos - rpos)) / sc.scale / sc.h
    return e_rho, e_lam, e_pos


@pytest.fixture(scope="module")
def ref64():
    sc = scenes.dam_break_8k()
    return sc, _oracle_one_step(sc, lf.LF_FLAG_DIFFUSE, fp64=True)


@pytest.mark.parametrize("flags", [0, lf.LF_FLAG_KAHAN, lf.LF_FLAG_FAST_MATH,
                                   lf.LF_FLAG_FAST_MATH | lf.LF_FLAG_KAHAN, lf.LF_FLAG_PACKED16,
                                   lf.LF_FLAG_PACKED16 | lf.LF_FLAG_FAST_MATH | lf.LF_FLAG_KAHAN])
def test_one_step_error_vs_fp64(ref64, flags):
    sc, ref = ref64
    got = _gpu_one_step(sc, flags | lf.LF_FLAG_DIFFUSE)
    e_rho, e_lam, e_pos = _errors(got, ref, sc)
    print("flags %#x: density %.3e lambda %.3e position %.3e" % (flags, e_rho, e_lam, e_pos))
    assert e_rho <= DENSITY_REL_TOL
    assert e_lam <= _lambda_tol(flags)
    assert e_pos <= POSITION_TOL


def test_fast_math_close_to_fp32_path(ref64):
    """FAST_MATH stays within a few FP32 roundings of the exact FP32 path."""
    sc, ref = ref64
    plain = _errors(_gpu_one_step(sc, lf.LF_FLAG_DIFFUSE), ref, sc)
    fast = _errors(_gpu_one_step(sc, lf.LF_FLAG_DIFFUSE | lf.LF_FLAG_FAST_MATH), ref, sc)
    for f, p in zip(fast, plain):
        assert f <= max(20 * p, 1e-6)


def test_diagnostics_fp64():
    sc = scenes.dam_break_8k()
    p = sc.params()
    with lf.Context(0) as c:
        c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        c.step(p, 3)
        d = c.diagnostics()
        d2 = c.diagnostics()
        st = c.download()
        rho = c.read(LF_ARR_DENSITY).astype(np.float64)
    assert d == d2, "FP64 reductions must be reproducible"
    fluid = st["type"] == lf.LF_FLUID
    m = st["mass"].astype(np.float64)
    assert d["fluid"] == int(fluid.sum())
    assert d["mass"] == pytest.approx(m[fluid].sum(), rel=1e-14)
    mom = (m[fluid, None] * st["velocity"][fluid].astype(np.float64)).sum(axis=0)
    for k in range(3):
        assert d["momentum"][k] == pytest.approx(mom[k], rel=1e-11, abs=1e-9)
    # rho is in sorted order, like the state arrays
    e = np.abs(rho[fluid] / np.float64(np.float32(6378.0)) - 1.0)
    assert d["density_error_mean"] == pytest.approx(e.mean(), rel=1e-12)
    assert d["density_error_max"] == pytest.approx(e.max(), rel=1e-12)
    assert d["passes"] == p.iterations


def _state(sc, p, steps=2):
    with lf.Context(0) as c:
        c.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        c.step(p, steps)
        st = c.download()
        d = c.diagnostics()
    return st, d


def test_convergence_met_at_first_pass_skips_deltas():
    sc = scenes.dam_break_8k()
    got, d = _state(sc, sc.params(density_tol=1e9))
    ref, _ = _state(sc, sc.params(iterations=0))
    assert d["passes"] == 1
    for k in ("position", "velocity", "order"):
        assert np.array_equal(got[k], ref[k]), k


def test_convergence_unreachable_changes_nothing():
    sc = scenes.dam_break_8k()
    got, d = _state(sc, sc.params(density_tol=1e-30))
    ref, _ = _state(sc, sc.params())
    assert d["passes"] == sc.iterations
    assert len(d["pass_error"]) == sc.iterations and all(e > 0 for e in d["pass_error"])
    for k in ("position", "velocity", "order"):
        assert np.array_equal(got[k], ref[k]), k


def test_convergence_per_scene_in_batch():
    sc = scenes.batch_scenes(3, lo=1000, hi=5000, seed=21)
    tols = [1e9, 0.0, 1e-30]
    params = [s.params(density_tol=t) for s, t in zip(sc, tols)]
    with lf.Context(0) as c:
        c.batch_upload(params, [dict(position=s.position, velocity=s.velocity, mass=s.mass)
                                for s in sc])
        c.batch_step(2)
        g = c.download()
        off = c.batch_info()
        diags = [c.diagnostics(s) for s in range(3)]
    assert diags[0]["passes"] == 1 and diags[1]["passes"] == sc[1].iterations
    assert diags[2]["passes"] == sc[2].iterations
    for s, scene in enumerate(sc):
        ref, _ = _state(scene, scene.params(iterations=0) if s == 0 else scene.params())
        part = g["position"][off[s]:off[s + 1]]
        order = g["order"][off[s]:off[s + 1]].astype(np.int64) - off[s]
        mine = np.empty_like(part)
        mine[order] = part
        theirs = np.empty_like(ref["position"])
        theirs[ref["order"].astype(np.int64)] = ref["position"]
        assert np.array_equal(mine, theirs), "scene %d" % s


def _report(tag, errs):
    print("%-34s density %.3e  lambda %.3e  position %.3e (of h)" % ((tag,) + tuple(errs)),
          flush=True)


MIXED = lf.LF_FLAG_FAST_MATH | lf.LF_FLAG_KAHAN
PACKED = MIXED | lf.LF_FLAG_PACKED16


def test_error_report_4m_one_step():
    """4,194,304-particle double dam break, one step from identical inputs
    (teacher forced): plain FP32 and mixed (FAST_MATH + KAHAN) against FP64."""
    sc = scenes.double_dam_break_4m()
    ref = _oracle_one_step(sc, lf.LF_FLAG_DIFFUSE, fp64=True)
    for tag, flags in (("4M one step, plain FP32", 0), ("4M one step, mixed", MIXED),
                       ("4M one step, mixed + packed16", PACKED)):
        errs = _errors(_gpu_one_step(sc, flags | lf.LF_FLAG_DIFFUSE), ref, sc)
        _report(tag, errs)
        assert errs[0] <= DENSITY_REL_TOL and errs[1] <= _lambda_tol(flags)
        assert errs[2] <= POSITION_TOL


def test_error_report_accumulated_20_steps():
    """8,000-particle dam break, 20 steps each side from the same start: the
    accumulated error of plain FP32 and mixed against the FP64 oracle."""
    sc = scenes.dam_break_8k()
    steps = 20
    ref = _oracle_state(sc, lf.LF_FLAG_DIFFUSE, True, steps)
    for tag, flags in (("8k 20 steps, plain FP32", 0), ("8k 20 steps, mixed", MIXED),
                       ("8k 20 steps, mixed + packed16", PACKED)):
        got = _gpu_state(sc, flags | lf.LF_FLAG_DIFFUSE, steps)
        errs = _errors(got, ref, sc)
        _report(tag, errs)
        # accumulated: a looser bar than one step (errors grow with the steps)
        assert errs[0] <= 1e-3 and errs[2] <= 1e-3


def test_packed16_needs_a_packed_grid():
    sc = scenes.small_block(n_side=6)
    with lf.Context(0) as c:
        c.upload(sc.position, sc.velocity, sc.mass)
        c.grid_build(sc.params())
        with pytest.raises(lf.LfError):
            c.lambda_pass(sc.params(flags=lf.LF_FLAG_PACKED16))
        c.grid_build(sc.params(flags=lf.LF_FLAG_PACKED16))
        c.lambda_pass(sc.params(flags=lf.LF_FLAG_PACKED16))
        c.delta_pass(sc.params(flags=lf.LF_FLAG_PACKED16))
        c.sync()
