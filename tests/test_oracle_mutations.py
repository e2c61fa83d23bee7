"""The oracle's pinning checks catch plausible mistakes (runs without a GPU).

Each case plants one mistake in a copy of the oracle source on the CPU -- a float compare
in place of the unsigned bit-pattern order, a dropped or misplaced nudge, a transposed
matrix, a sign slip in a derivative coefficient or solver term, a dropped weight, a short
or inverted bisection -- builds it into a temporary library, and runs the geometry pins
against it through SCANLINE_ORACLE_LIB.  They must fail for every case.  Each planted
snippet must occur exactly once in the real source, so an edit of the oracle that moves
one of them is noticed here.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SRC = os.path.join(ORACLE, "scanline_oracle.cpp")
PINS = os.path.join(ROOT, "tests", "test_oracle_geometry_pins.py")

MUTATIONS = {
    "orientation by float compare": (
        "if (iyf < iyl || (iyf == iyl && ixf < ixl)) {",
        "if (s.vy[p] < s.vy[p + vn - 1] || (s.vy[p] == s.vy[p + vn - 1] && s.vx[p] < s.vx[p + vn - 1])) {"),
    "nudge dropped": ("if (std::floor(o[0]) == o[0]) o[0] = o[0] - 1.f / 256.f;", ""),
    "nudge on y": ("if (std::floor(o[0]) == o[0]) o[0] = o[0] - 1.f / 256.f;",
                   "if (std::floor(o[1]) == o[1]) o[1] = o[1] - 1.f / 256.f;"),
    "transposed matrix": ("for (int j = 0; j < 4; ++j) rows[i][j] = m[j * 4 + i];",
                          "for (int j = 0; j < 4; ++j) rows[i][j] = m[i * 4 + j];"),
    "depth clamp dropped": ("if (o[2] < cutoff) o[2] = cutoff;", ""),
    "region slots swapped": ("case 0x12: slot0 = 4; break;", "case 0x12: slot0 = 3; break;"),
    "cubic split: sign of b": ("2.f * ((v[0] - v[1]) + (v[2] - v[1])),\n                       v[1] - v[0]",
                               "-2.f * ((v[0] - v[1]) + (v[2] - v[1])),\n                       v[1] - v[0]"),
    "cubic split: transposed a": ("solve_quad((v[3] - v[0]) + 3.f * (v[1] - v[2])",
                                  "solve_quad((v[3] - v[0]) + 3.f * (v[2] - v[1])"),
    "rational split: sign of b": ("-(2.f * pw01 + pw20)", "(2.f * pw01 + pw20)"),
    "quadratic split: sign": ("float t = (v[0] - v[1]) / b;", "float t = (v[1] - v[0]) / b;"),
    "quadratic solve: sign of b": ("float b = -2.f * iA + 2.f * iB;", "float b = 2.f * iA - 2.f * iB;"),
    "rational solve: weight dropped": ("(v[1] - c) * cd.w, v[2] - c", "(v[1] - c), v[2] - c"),
    "bisection: 20 steps": ("for (int it = 0; it < 24; ++it) {", "for (int it = 0; it < 20; ++it) {"),
    "bisection: inverted bracket": (
        "if ((int32_t)(fbits(vtm - c) ^ fbits(vt0 - c)) >= 0) { a0 = tm; vt0 = vtm; }",
        "if ((int32_t)(fbits(vtm - c) ^ fbits(vt0 - c)) < 0) { a0 = tm; vt0 = vtm; }"),
}


def _build(args):
    name, (old, new), src, out = args
    mutated = os.path.join(os.path.dirname(out), "mutated.cpp")
    with open(mutated, "w") as f:
        f.write(src.replace(old, new))
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared",
                    "-I", ORACLE, "-o", out, mutated], check=True)
    return out


@pytest.fixture(scope="module")
def mutated_libs(tmp_path_factory):
    src = open(SRC).read()
    jobs = []
    for k, (name, (old, new)) in enumerate(MUTATIONS.items()):
        assert src.count(old) == 1, "planted snippet not found once in the oracle: %s" % name
        d = tmp_path_factory.mktemp("mut%02d" % k)
        jobs.append((name, (old, new), src, str(d / "liboracle.so")))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        outs = list(ex.map(_build, jobs))
    return dict(zip(MUTATIONS, outs))


def _run_pins(lib):
    env = dict(os.environ, SCANLINE_ORACLE_LIB=lib)
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", PINS],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    return p.returncode, p.stdout


def test_every_planted_mistake_fails_a_pin(mutated_libs):
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = dict(zip(mutated_libs, ex.map(_run_pins, mutated_libs.values())))
    missed = [name for name, (rc, _) in results.items() if rc != 1]
    assert not missed, {n: results[n][1][-400:] for n in missed}


def test_the_real_oracle_passes_the_pins():
    rc, out = _run_pins(os.path.join(ORACLE, "build", "liboracle.so"))
    assert rc == 0, out[-2000:]
