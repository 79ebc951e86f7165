"""The strided-NCC and warp pins catch plausible mistakes.

Each case plants one mistake in a copy of the oracle (oracle/pm_oracle.cc),
builds it on the CPU with the oracle's own flags, and runs
tests/test_oracle_strided.py against it (PM_ORACLE_LIB): the pins must
fail.  The unmodified oracle passes them (that file itself)."""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "pm_oracle.cc")
FLAGS = ["-O1", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-w"]

# (name, text in pm_oracle.cc, planted mistake)
MUTANTS = [
    ("strided weight without the stride",
     "const real w = rexp(fdx * fdx * c.spatial_factor + fdy * fdy * c.spatial_factor +",
     "const real w = rexp(static_cast<real>(ix * ix) * c.spatial_factor + fdy * fdy * c.spatial_factor +"),
    ("strided source x/y Jacobian terms swapped",
     "const real sx = rfma(jxx, fdx, bx);\n      const real sy = rfma(jxy, fdx, by);\n      const real src_val",
     "const real sx = rfma(jxy, fdx, bx);\n      const real sy = rfma(jxx, fdx, by);\n      const real src_val"),
    ("strided reference row off by one",
     "const real ref_val = c.ref.texel(static_cast<int>(rx), static_cast<int>(ry));",
     "const real ref_val = c.ref.texel(static_cast<int>(rx), static_cast<int>(ry) + 1);"),
    ("strided row offset not scaled by the stride",
     "const real bx = rfma(jyx, fdy, scx);\n    const real by = rfma(jyy, fdy, scy);",
     "const real bx = rfma(jyx, static_cast<real>(iy), scx);\n    const real by = rfma(jyy, static_cast<real>(iy), scy);"),
    ("strided covariance without r",
     "sum_rs = rfma(w * r, sv, sum_rs);",
     "sum_rs = rfma(w, sv, sum_rs);"),
    ("warp Jacobian transposed",
     "  w->jxy = pdx.y / pdx.z - w->scy;\n  w->jyx = pdy.x / pdy.z - w->scx;",
     "  w->jxy = pdy.x / pdy.z - w->scx;\n  w->jyx = pdx.y / pdx.z - w->scy;"),
    ("warp forward difference taken backwards",
     "  w->jxx = pdx.x / pdx.z - w->scx;",
     "  w->jxx = w->scx - pdx.x / pdx.z;"),
]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    src = open(SRC).read()
    out = tmp_path_factory.mktemp("mutants")
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.fail("g++ is needed to build the oracle mutants")

    def build(i):
        if i < 0:  # control: the oracle as it is, built the same way
            name, old, new = "control", "", ""
        else:
            name, old, new = MUTANTS[i]
            assert src.count(old) == 1, name
        cc = out / f"m{i}.cc"
        cc.write_text(src.replace(old, new) if old else src)
        so = out / f"m{i}.so"
        subprocess.run([cxx, *FLAGS, str(cc), "-o", str(so)], check=True)
        return str(so)

    with ThreadPoolExecutor(max_workers=4) as pool:
        return list(pool.map(build, range(-1, len(MUTANTS))))


def _pins_pass(lib: str) -> bool:
    env = dict(os.environ, PM_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_strided.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode == 0


def test_every_planted_mistake_fails_the_pins(mutant_libs):
    with ThreadPoolExecutor(max_workers=4) as pool:
        passed = list(pool.map(_pins_pass, mutant_libs))
    assert passed[0], "the unmodified oracle, built the same way, must pass the pins"
    survivors = [MUTANTS[i][0] for i, ok in enumerate(passed[1:]) if ok]
    assert not survivors, survivors
