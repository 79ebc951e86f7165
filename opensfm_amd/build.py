"""Build the native parts in-tree.

  opensfm_amd/libopensfm_pm.so   HIP kernels + C ABI (hipcc, gfx950)
  opensfm_amd/pydense*.so        pybind11 module over the C ABI (g++)

Run ``python -m opensfm_amd.build``; the package never builds on import.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libopensfm_pm.so")
ARCH = os.environ.get("PM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # numerics contract: no implicit contraction, IEEE divide / sqrt
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    # no packed-FP32 VALU ops anywhere
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# HIP translation units of the library, compiled in parallel.
SOURCES = ("pm_kernels.hip", "pm_cost_f32.hip", "pm_cost_f16.hip", "pm_refine.hip", "pm_clean.hip", "pm_prior.hip",
           "pm_api.hip")


def pydense_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "pydense" + suffix)


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_hip(out: str = LIB, defines=(), sources=None, extra=()) -> str:
    """Compile the HIP sources (in parallel) and link ``out``.

    ``defines``: extra -D flags (tuning knobs); ``sources``: override the
    source paths (default: SOURCES under csrc/);
    ``extra``: further compiler flags.
    """
    srcs = sources or [os.path.join(CSRC, n) for n in SOURCES]
    objs = []
    procs = []
    for src in srcs:
        obj = out + "." + os.path.basename(src).replace(".hip", ".o")
        cmd = ["hipcc", *HIP_FLAGS, *extra, *[f"-D{d}" for d in defines], f"-I{CSRC}", "-c", src, "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), src))
        objs.append(obj)
    for p, src in procs:
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed on {src}")
    tmp = out + ".tmp"
    _run(["hipcc", "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


def build_pydense() -> str:
    import pybind11

    out = pydense_path()
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
           f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(CSRC, "pydense.cpp"), os.path.join(CSRC, "lz4_frame.cpp"), "-o", tmp,
           f"-L{HERE}", "-lopensfm_pm", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_hip()
    build_pydense()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "hip": build_hip, "pydense": build_pydense}[what]()
