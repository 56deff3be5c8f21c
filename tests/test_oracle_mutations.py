"""Planted mistakes in the oracle must be caught.

Each case below is a plausible misreading of the reference's constraint
arithmetic (a wrong exponent, sign, factor or index in sph_lambda /
sph_delta, oclsph_kernel.h:305-384; in the wells of advectAndCopy,
oclsph.hpp:306-317; in the XSPH coefficient). The test writes a copy of
oracle/pbf_oracle.hpp with the mistake planted, builds it beside a copy of
oracle/oracle_capi.cpp into a temporary library, and runs the pins of
tests/test_oracle_pins.py against that library: at least one pin must fail
for every mistake, and the unmodified build must pass them all. The planted
copies live only in a temporary directory, on the CPU.
"""
import concurrent.futures
import os
import shutil
import subprocess

import pytest

from tests import test_oracle_pins as pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# (name, text in pbf_oracle.hpp, replacement)
MUTATIONS = [
    ("scorr_exponent_2",
     "const F corr = -c.corrK * ((q * q) * (q * q));",
     "const F corr = -c.corrK * (q * q);"),
    ("scorr_sign",
     "const F corr = -c.corrK * ((q * q) * (q * q));",
     "const F corr = c.corrK * ((q * q) * (q * q));"),
    ("lambda_gradient_x10",
     "for (int k = 0; k < 3; ++k) g[k] = std::fma(gv[k], c.rhoRecip, g[k]);",
     "for (int k = 0; k < 3; ++k) g[k] = std::fma(gv[k] * F(10), c.rhoRecip, g[k]);"),
    ("delta_factor_sign",
     "const F factor = ((la + lambda[b]) + corr) * c.rhoRecip;",
     "const F factor = -(((la + lambda[b]) + corr) * c.rhoRecip);"),
    ("delta_lambda_b_index",
     "const F factor = ((la + lambda[b]) + corr) * c.rhoRecip;",
     "const F factor = ((la + la) + corr) * c.rhoRecip;"),
    ("delta_clamp_dropped",
     "const F q = std::min(c.maxB[k], std::max(c.minB[k], pos[k])); /* :371 */",
     "const F q = pos[k];"),
    ("density_mass_of_b",
     "r = std::fma(ma, w, r);",
     "r = std::fma(ps[4 * b + 3], w, r);"),
    ("cfm_dropped",
     "lambda[a] = -C1 / (norm2 + c.cfm);",
     "lambda[a] = -C1 / (norm2 + F(1));"),
    ("xsph_coefficient_x2",
     "for (int k = 0; k < 3; ++k) out[3 * a + k] = std::fma(c.xsphRho, acc[k], va[k]);",
     "for (int k = 0; k < 3; ++k) out[3 * a + k] = std::fma(F(2) * c.xsphRho, acc[k], va[k]);"),
    ("xsph_difference_sign",
     "acc[k] = std::fma(mw, vs[3 * b + k] - va[k], acc[k]);",
     "acc[k] = std::fma(mw, va[k] - vs[3 * b + k], acc[k]);"),
    ("well_force_sign",
     "F f = ((d[k] * inv) * F(w.force)) * mass[i];",
     "F f = -(((d[k] * inv) * F(w.force)) * mass[i]);"),
    ("well_clamp_dropped",
     "f = std::min(std::max(f, F(-10)), F(10));",
     "f = f;"),
]


def _build(tmp, name, old, new):
    d = os.path.join(tmp, name)
    os.makedirs(d)
    src = open(os.path.join(ORACLE, "pbf_oracle.hpp")).read()
    if old is not None:
        assert src.count(old) == 1, "mutation %s: the planted text no longer matches" % name
        src = src.replace(old, new)
    open(os.path.join(d, "pbf_oracle.hpp"), "w").write(src)
    shutil.copy(os.path.join(ORACLE, "oracle_capi.cpp"), d)
    out = os.path.join(d, "liboracle_%s.so" % name)
    subprocess.run(["g++", "-O1", "-fopenmp", "-ffp-contract=off", "-std=c++17", "-fPIC", "-shared",
                    "-Wl,-Bsymbolic", "-I" + d, "-I" + os.path.join(ROOT, "include"),
                    os.path.join(d, "oracle_capi.cpp"), "-o", out], check=True,
                   capture_output=True)
    return out


def _failing_pins(lib_path):
    failed = []
    for pin in pins.PINS:
        try:
            pin(lib_path)
        except AssertionError:
            failed.append(pin.__name__)
    return failed


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_planted_mistakes_fail_a_pin(tmp_path):
    cases = [("unmodified", None, None)] + MUTATIONS
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        libs = list(ex.map(lambda c: _build(str(tmp_path), *c), cases))
    assert _failing_pins(libs[0]) == [], "the unmodified oracle must pass every pin"
    caught = {name: _failing_pins(lib) for (name, _, _), lib in zip(cases[1:], libs[1:])}
    missed = [name for name, failed in caught.items() if not failed]
    assert not missed, "planted mistakes no pin caught: %s" % missed
