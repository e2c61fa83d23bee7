"""The chord-animation oracle's pins catch plausible mistakes (runs without a GPU).

Each case plants one mistake in a copy of oracle/chord_oracle.cpp on the CPU -- a handle
pointing the wrong way or sized on the wrong radius, a misplaced middle angle, (cos, sin)
for (sin, cos), a sign slip in the series or a quadrant, a dropped reduction term, a
flattened ease, a wrong chord radius, a link end without its position, the colour of the
smaller end, the keyframes swapped, an inverted orientation test -- links it with the rest
of the oracle into a temporary library, and runs tests/test_chord_pins.py against it
through SCANLINE_ORACLE_LIB.  Every case must fail a pin.  Each planted snippet must occur
exactly once in the real source.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SRC = os.path.join(ORACLE, "chord_oracle.cpp")
PINS = os.path.join(ROOT, "tests", "test_chord_pins.py")
FLAGS = ["-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math"]

MUTATIONS = {
    "handle sign": ("p[1] = sub(p[0], mul(n0, len));", "p[1] = add(p[0], mul(n0, len));"),
    "handle on the inner radius": ("sin_ref((t1 - t0) * .5f) * .353f * R;", "sin_ref((t1 - t0) * .5f) * .353f * kSmallR;"),
    "middle angle": ("(float)(t0 * 0.5 + t1 * 0.5)", "(float)(t0 * 0.5 + t1 * 0.25)"),
    "cos for sin": ("sincos_ref(ts[i], &d.x, &d.y);", "sincos_ref(ts[i], &d.y, &d.x);"),
    "series sign": ("0.0001984126984126984, -0.008333333333333333,", "0.0001984126984126984, 0.008333333333333333,"),
    "quadrant sign": ("case 1: s = cr; c = -sr; break;", "case 1: s = cr; c = sr; break;"),
    "reduction tail dropped": ("const double r = (x - k * head) - k * tail;", "const double r = x - k * head;"),
    "ease without its factor": ("t < 0.5f ? 2 * t * t :", "t < 0.5f ? t * t :"),
    "chord radius": ("kSmallR / (2 * sin_ref(th_s) + 1.f)", "kSmallR / (sin_ref(th_s) + 1.f)"),
    "link end without position": ("(lpos0[i] + lsz0[i]) / ns0", "lsz0[i] / ns0"),
    "colour of the smaller end": ("lsz0[i] > lsz1[i] ? n0 : n1", "lsz0[i] > lsz1[i] ? n1 : n0"),
    "keyframes swapped": ("at0[i] = node_t0[p] * (1 - t) + node_t0[n] * t;",
                          "at0[i] = node_t0[n] * (1 - t) + node_t0[p] * t;"),
    "orientation inverted": ("if (!(iyf < iyl || (iyf == iyl && ixf < ixl))) {",
                             "if (iyf < iyl || (iyf == iyl && ixf < ixl)) {"),
}


@pytest.fixture(scope="module")
def mutated_libs(tmp_path_factory):
    src = open(SRC).read()
    base = tmp_path_factory.mktemp("chord_base")
    objs = []
    for name in ("scanline_oracle.cpp", "diff_oracle.cpp"):
        o = str(base / (name + ".o"))
        subprocess.run(["g++", *FLAGS, "-c", "-I", ORACLE, "-o", o, os.path.join(ORACLE, name)], check=True)
        objs.append(o)
    jobs = []
    for k, (name, (old, new)) in enumerate(MUTATIONS.items()):
        assert src.count(old) == 1, "planted snippet not found once in the oracle: %s" % name
        d = tmp_path_factory.mktemp("chord_mut%02d" % k)
        with open(d / "chord.cpp", "w") as f:
            f.write(src.replace(old, new))
        jobs.append((str(d / "chord.cpp"), str(d / "liboracle.so")))

    def build(job):
        cpp, out = job
        subprocess.run(["g++", *FLAGS, "-shared", "-o", out, cpp, *objs], check=True)
        return out

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        outs = list(ex.map(build, jobs))
    return dict(zip(MUTATIONS, outs))


def _run_pins(lib):
    env = dict(os.environ, SCANLINE_ORACLE_LIB=lib)
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", PINS],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    return p.returncode, p.stdout


def test_every_planted_mistake_fails_a_chord_pin(mutated_libs):
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = dict(zip(mutated_libs, ex.map(_run_pins, mutated_libs.values())))
    missed = [name for name, (rc, _) in results.items() if rc != 1]
    assert not missed, {n: results[n][1][-400:] for n in missed}


def test_the_real_chord_oracle_passes_the_pins():
    rc, out = _run_pins(os.path.join(ORACLE, "build", "liboracle.so"))
    assert rc == 0, out[-2000:]
