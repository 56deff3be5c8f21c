"""Pins of the oracle's constraint arithmetic (oracle/pbf_oracle.hpp) that the
lattice and symmetry checks of test_oracle.py leave open: the lambda pass's
gradient sum, the delta pass's s_corr term, lambda_a + lambda_b factor and
clamp, the gravity wells of the predict step, and the XSPH coefficient.

Each check restates the reference formula without cells, in FP64 numpy over
every pair (or as a closed form with hand-derived numbers), on inputs where
the term in question is large: a jittered asymmetric cluster with boundary
particles (so the gradient sum is far from zero), random lambdas of both
signs, particles pressed against the box. A wrong coefficient, exponent,
sign or index moves the compared values by far more than the tolerance;
tests/test_oracle_mutations.py plants such mistakes in the oracle and checks
that these functions reject every one.

Reference formulas (paths relative to the reference root):
  poly6 W(r) = 315/(64 pi h^9) (h^2 - r^2)^3 for r <= h   oclsph_kernel.h:50-56
  spiky grad(a,b) = (p_a - p_b) (-45/(pi h^6)) (h - r)^2 / r
                    for 1e-8 <= r <= h                       oclsph_kernel.h:61-71
  lambda_a = -(rho_a/rho0 - 1) / (|sum_b grad(a,b)/rho0|^2 + 600),
  rho_a = sum_b m_a W(r_ab) over every candidate, obstacles included;
  obstacles: lambda = 0                                      oclsph_kernel.h:305-332
  dp_a = sum_b grad(a,b) (lambda_a + lambda_b - 1e-4 (W(r)/W(0.3h))^4) / rho0;
  p*_a <- clamp((p*_a + dp_a) scale, minBound, maxBound) / scale;
  obstacles unchanged                                        oclsph_kernel.h:335-384
  F = m g + sum_wells clamp(rhat f m / d^2, -10, 10) for d < 75;
  v += F dt; p* = v dt + x / scale                          oclsph.hpp:289-325
  XSPH (not in the reference): v_a += (c/rho0) sum_b m_b (v_b - v_a) W(r_ab),
  obstacles with zero weight.
"""
import itertools
import math

import numpy as np
import pytest

import libfluid_amd as lf
import oracle
from libfluid_amd._abi import LF_ARR_DENSITY, LF_ARR_LAMBDA, LF_ARR_PSTAR, LF_ARR_VELOCITY

RHO0 = 6378.0
CFM = 600.0
CORR_K = float(np.float32(1e-4))   # lf_pbf_params holds the constants as FP32
CORR_DQ = float(np.float32(0.3))
OBSTACLE = lf.LF_OBSTACLE


# ---- the reference formulas, over every pair ------------------------------------

def poly6(r, h):
    return np.where(r <= h, 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3, 0.0)


def spiky_grad(d, r, h):
    """d = p_a - p_b (..., 3), r = |d|."""
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where((r >= 1e-8) & (r <= h), -45.0 / (math.pi * h ** 6) * (h - r) ** 2 / r, 0.0)
    return d * s[..., None]


def pairs(ps):
    d = ps[:, None, :] - ps[None, :, :]
    return d, np.sqrt((d * d).sum(-1))


def lambda_brute(ps, mass, types, h):
    d, r = pairs(ps)
    rho = mass * poly6(r, h).sum(1)
    g = (spiky_grad(d, r, h) / RHO0).sum(1)
    lam = -(rho / RHO0 - 1.0) / ((g * g).sum(-1) + CFM)
    lam[types == OBSTACLE] = 0.0
    rho[types == OBSTACLE] = 0.0
    return lam, rho, g


def delta_brute(ps, lam, types, h, lo, hi, scale=1.0, corr_k=CORR_K):
    d, r = pairs(ps)
    corr = -corr_k * (poly6(r, h) / poly6(np.float64(CORR_DQ * h), h)) ** 4
    factor = (lam[:, None] + lam[None, :] + corr) / RHO0
    dp = (spiky_grad(d, r, h) * factor[..., None]).sum(1)
    out = np.clip((ps + dp) * scale, lo, hi) / scale
    out[types == OBSTACLE] = ps[types == OBSTACLE]
    return out, dp


def xsph_brute(ps, vel, mass, types, h, c):
    d, r = pairs(ps)
    w = poly6(r, h) * np.where(types == OBSTACLE, 0.0, mass)[None, :]
    dv = (w[..., None] * (vel[None, :, :] - vel[:, None, :])).sum(1)
    out = vel + (c / RHO0) * dv
    out[types == OBSTACLE] = vel[types == OBSTACLE]
    return out


# ---- inputs -----------------------------------------------------------------------

H = 0.125  # exact in binary; the cell size of the oracle's grid


def cluster(seed, centre=(0.5, 0.5, 0.5), axes=(0.33, 0.25, 0.19), obstacles=0.1):
    """A jittered, anisotropic blob of ~270 particles: a lattice of spacing h/2
    (33 lattice points within h) inside an ellipsoid, jittered by 0.3 spacing,
    in random order, some particles obstacles, masses varying 0.5-2. Its surface
    particles have large one-sided gradients."""
    rng = np.random.default_rng(seed)
    s = H / 2
    ax = np.arange(0.0, 1.0, s) + s / 3
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    q = (g - np.asarray(centre)) / np.asarray(axes)
    g = g[(q * q).sum(1) <= 1.0]
    g = g + rng.uniform(-0.3 * s, 0.3 * s, g.shape)
    g = g[rng.permutation(len(g))]
    types = np.where(rng.random(len(g)) < obstacles, OBSTACLE, lf.LF_FLUID).astype(np.int32)
    mass = rng.uniform(0.5, 2.0, len(g))
    return g, mass, types


def sorted_oracle(pos, mass, types, lib_path=None, **kw):
    """An FP64 oracle after its grid build (zero velocity and gravity, so
    p* = x), with the sorted p*, masses and types."""
    kw.setdefault("gravity", (0.0, 0.0, 0.0))
    kw.setdefault("max_bound", (1.0, 1.0, 1.0))
    p = lf.pbf_params(H, **kw)
    o = oracle.Pbf(p, fp64=True, lib_path=lib_path)
    o.load(pos, np.zeros_like(pos), mass, types)
    o.grid_build()
    ps = o.get(LF_ARR_PSTAR)
    return o, p, ps[:, :3].copy(), ps[:, 3].copy(), o.types()


# ---- the pins, each a function of the oracle build under test -------------------------

def check_lambda_cluster(lib_path=None):
    pos, mass, types = cluster(21)
    o, _, ps, ms, ts = sorted_oracle(pos, mass, types, lib_path)
    o.lambda_pass()
    lam, rho, g = lambda_brute(ps, ms, ts, H)
    fluid = ts != OBSTACLE
    # the inputs exercise what is pinned: gradient sums that change lambda by
    # tens of percent against the CFM term alone, densities on both sides of rho0
    g2 = (g * g).sum(1)[fluid]
    assert g2.max() > 0.1 * CFM and np.median(g2) > 1.0
    assert (rho[fluid] > RHO0).any() and (rho[fluid] < RHO0).any()
    got_rho, got_lam = o.get(LF_ARR_DENSITY), o.get(LF_ARR_LAMBDA)
    assert np.allclose(got_rho, rho, rtol=1e-12, atol=1e-9)
    assert np.allclose(got_lam, lam, rtol=1e-10, atol=1e-15)


def check_delta_cluster(lib_path=None):
    """Teacher-forced lambdas (random, both signs) so the delta pass is pinned
    on its own; part of the blob sits against the box's upper x and y bounds so
    the clamp engages."""
    pos, mass, types = cluster(22, centre=(0.6, 0.62, 0.5))
    lo_b, hi_b = np.array([0.0, 0.0, 0.0]), np.array([0.84375, 0.859375, 1.0])  # FP32-exact
    o, _, ps, ms, ts = sorted_oracle(pos, mass, types, lib_path, max_bound=tuple(hi_b))
    rng = np.random.default_rng(23)
    lam = rng.uniform(-2e-4, 6e-4, len(ps))
    lam[ts == OBSTACLE] = 0.0
    o.set(LF_ARR_LAMBDA, lam)
    o.delta_pass()
    got = o.get(LF_ARR_PSTAR)[:, :3]
    want, dp = delta_brute(ps, lam, ts, H, lo_b, hi_b)
    # sensitivity: the s_corr term alone moves particles by far more than the
    # tolerance, and the clamp holds some of them on the bounds
    _, dp0 = delta_brute(ps, lam, ts, H, lo_b, hi_b, corr_k=0.0)
    fluid = ts != OBSTACLE
    assert np.abs(dp - dp0)[fluid].max() > 1e-7
    assert np.abs(dp[fluid]).max() > 1e-5
    clamped = np.any((ps + dp > hi_b) | (ps + dp < lo_b), axis=1) & fluid
    assert clamped.sum() >= 5
    assert np.allclose(got, want, rtol=0, atol=1e-13)


def check_two_particles(lib_path=None):
    """Two fluid particles of mass 1 at distance h/2 along x, hand-derived:
    W(0) = 315/(64 pi h^3), W(h/2) = 315 (3/4)^3 / (64 pi h^3),
    |grad| = 45/(pi h^6) (h/2)^2 / (h/2) * (h/2) = 45/(4 pi h^4) along +x for the
    left particle; rho = W(0) + W(h/2); lambda = -(rho/rho0 - 1)/(g^2 + 600),
    g = |grad|/rho0; the left particle's delta is
    |grad| (2 lambda - 1e-4 ((3/4)^3/(0.91)^3)^4) / rho0 along +x."""
    x0 = 0.5
    pos = np.array([[x0, 0.5, 0.5], [x0 + H / 2, 0.5, 0.5]])
    o, _, ps, _, _ = sorted_oracle(pos, np.ones(2), np.zeros(2, np.int32), lib_path)
    assert ps[0, 0] < ps[1, 0]
    w0 = 315.0 / (64.0 * math.pi * H ** 3)
    wh = w0 * 0.75 ** 3
    grad = 45.0 / (4.0 * math.pi * H ** 4)
    rho = w0 + wh
    g = grad / RHO0
    lam = -(rho / RHO0 - 1.0) / (g * g + CFM)
    o.lambda_pass()
    assert np.allclose(o.get(LF_ARR_DENSITY), [rho, rho], rtol=1e-13)
    assert np.allclose(o.get(LF_ARR_LAMBDA), [lam, lam], rtol=1e-12)
    wdq = w0 * (1.0 - CORR_DQ ** 2) ** 3
    corr = -CORR_K * (wh / wdq) ** 4
    dx = grad * (2.0 * lam + corr) / RHO0
    o.delta_pass()
    got = o.get(LF_ARR_PSTAR)[:, :3]
    assert abs(dx) > 1e-3
    assert np.allclose(got[0], [x0 + dx, 0.5, 0.5], rtol=0, atol=1e-14)
    assert np.allclose(got[1], [x0 + H / 2 - dx, 0.5, 0.5], rtol=0, atol=1e-14)


def check_boundary_lattice(lib_path=None):
    """A particle on the x = min face of an unjittered lattice (spacing h/2,
    interior in y and z): rho = W-sum over the lattice points k (k_x >= 0,
    |k| <= 2 spacings), and by the y/z symmetry g = (g_x, 0, 0) with
    g_x = sum_k (-k_x s)(-45/(pi h^6))(h - |k| s)^2 / (|k| s) / rho0."""
    s = H / 2
    n_side = 10
    ii = np.arange(n_side, dtype=np.float64)
    pos = 0.3 + s * np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    o, _, ps, _, _ = sorted_oracle(pos, np.ones(len(pos)), np.zeros(len(pos), np.int32), lib_path)
    o.lambda_pass()
    a = int(np.argmin(np.abs(ps - np.array([0.3, 0.3 + 5 * s, 0.3 + 5 * s])).sum(1)))
    assert np.allclose(ps[a], [0.3, 0.3 + 5 * s, 0.3 + 5 * s])
    poly = 315.0 / (64.0 * math.pi * H ** 9)
    rho = gx = 0.0
    for k in itertools.product(range(0, 3), range(-2, 3), range(-2, 3)):
        q = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
        if q > 4:
            continue
        r = math.sqrt(q) * s
        rho += poly * (H * H - r * r) ** 3
        if q:
            gx += (-k[0] * s) * (-45.0 / (math.pi * H ** 6)) * (H - r) ** 2 / r / RHO0
    lam = -(rho / RHO0 - 1.0) / (gx * gx + CFM)
    assert gx * gx > 10.0  # a few percent of the CFM term: a wrong gradient factor shows
    assert math.isclose(o.get(LF_ARR_DENSITY)[a], rho, rel_tol=1e-13)
    assert math.isclose(o.get(LF_ARR_LAMBDA)[a], lam, rel_tol=1e-12)


def check_wells(lib_path=None):
    """One isolated particle, m = 2 at the origin, v0 = (0.5, 0, 0.25),
    g = (0, -1, 0), dt = 1/8, scale 1 and three wells: (3,4,0) force 50 at
    distance 5 adds rhat 50*2/25 = (2.4, 3.2, 0); (0,0,-2) force 100 at distance
    2 adds (0, 0, -50) clamped to (0, 0, -10); (80,0,0) lies beyond 75 and adds
    nothing. F = (0,-2,0) + (2.4,3.2,0) + (0,0,-10) = (2.4, 1.2, -10),
    v = v0 + F dt = (0.8, 0.15, -1.0), p* = v dt = (0.1, 0.01875, -0.125)."""
    p = lf.pbf_params(H, dt=0.125, gravity=(0.0, -1.0, 0.0), min_bound=(-1.0, -1.0, -1.0),
                      max_bound=(1.0, 1.0, 1.0))
    o = oracle.Pbf(p, fp64=True, lib_path=lib_path)
    o.set_wells([lf.WellSpec(centre=(3.0, 4.0, 0.0), force=50.0),
                 lf.WellSpec(centre=(0.0, 0.0, -2.0), force=100.0),
                 lf.WellSpec(centre=(80.0, 0.0, 0.0), force=1e6)])
    o.load(np.zeros((1, 3)), np.array([[0.5, 0.0, 0.25]]), np.array([2.0]))
    o.grid_build()
    assert np.allclose(o.get(LF_ARR_PSTAR)[0, :3], [0.1, 0.01875, -0.125], rtol=0, atol=1e-15)


def check_xsph_two_particles(lib_path=None):
    """Two fluid particles at distance h/2, masses 1 and 2, c = 1/64:
    dv_a = (c/rho0) m_b (v_b - v_a) W(h/2) with W(h/2) = 315 (3/4)^3 / (64 pi h^3)."""
    c = 1.0 / 64.0
    pos = np.array([[0.5, 0.5, 0.5], [0.5, 0.5 + H / 2, 0.5]])
    p = lf.pbf_params(H, gravity=(0.0, 0.0, 0.0), xsph_c=c, iterations=0)
    o = oracle.Pbf(p, fp64=True, lib_path=lib_path)
    o.load(pos, np.zeros((2, 3)), np.array([1.0, 2.0]))
    o.grid_build()
    o.finalise()
    m = o.get(LF_ARR_PSTAR)[:, 3]
    vel = np.array([[0.25, -0.5, 1.0], [-0.75, 0.5, 0.125]])
    o.set(LF_ARR_VELOCITY, vel)
    o.xsph()
    wh = 315.0 * 0.75 ** 3 / (64.0 * math.pi * H ** 3)
    want = vel + (c / RHO0) * wh * m[::-1, None] * (vel[::-1] - vel)
    assert np.abs(want - vel).max() > 1e-4
    assert np.allclose(o.get(LF_ARR_VELOCITY), want, rtol=0, atol=1e-15)


def check_xsph_cluster(lib_path=None):
    pos, mass, types = cluster(24)
    c = 0.01
    p = lf.pbf_params(H, gravity=(0.0, 0.0, 0.0), xsph_c=c, iterations=0)
    o = oracle.Pbf(p, fp64=True, lib_path=lib_path)
    o.load(pos, np.zeros_like(pos), mass, types)
    o.grid_build()
    o.finalise()
    ps = o.get(LF_ARR_PSTAR)
    ts = o.types()
    vel = np.random.default_rng(25).uniform(-1.0, 1.0, (len(ps), 3))
    o.set(LF_ARR_VELOCITY, vel)
    o.xsph()
    want = xsph_brute(ps[:, :3], vel, ps[:, 3], ts, H, float(np.float32(c)))
    assert np.abs(want - vel).max() > 1e-3
    assert np.allclose(o.get(LF_ARR_VELOCITY), want, rtol=0, atol=1e-13)


PINS = [check_lambda_cluster, check_delta_cluster, check_two_particles, check_boundary_lattice,
        check_wells, check_xsph_two_particles, check_xsph_cluster]


@pytest.mark.parametrize("pin", PINS, ids=[f.__name__ for f in PINS])
def test_oracle_pin(pin):
    pin()
