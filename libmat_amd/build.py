"""Builds the gfx950 shared libraries libmatgpu_f32.so / libmatgpu_f64.so.

Every HIP source is compiled once to an object under libmat_amd/_build/;
the drop-in layer (capi_dropin.cpp, mat_host.cpp) is compiled once per element type
(MAT_DOUBLE_PRECISION for the double flavour, as libmat selects
mat_elem_t); both libraries link the same kernel objects.  Usage:
    python -m libmat_amd.build   (or libmat_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["gemm.hip", "lowp.hip", "panel.hip", "trsm.hip", "chol.hip", "lu.hip",
               "solve.hip", "trsv.hip", "rbt.hip", "batched.hip", "tslu.hip", "qr.hip"]
CPP_SOURCES = ["runtime.cpp", "capi_ext.cpp"]
# compiled once per element type (MAT_DOUBLE_PRECISION for the double flavour)
PER_TYPE = ["capi_dropin.cpp", "mat_host.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}"]
# Packed fp32 VALU ops stay disabled in all device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
    return err


def _obj(src, tag=""):
    base = os.path.splitext(os.path.basename(src))[0]
    return os.path.join(BUILD, f"{base}{tag}.o")


def build(verbose: bool = False, jobs: int | None = None) -> list[str]:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    tasks = []
    for s in HIP_SOURCES:
        tasks.append([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in CPP_SOURCES:
        tasks.append([HIPCC, *COMMON, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for tag, defs in (("_f32", []), ("_f64", ["-DMAT_DOUBLE_PRECISION"])):
        for src in PER_TYPE:
            tasks.append([HIPCC, *COMMON, *defs, "-c", os.path.join(CSRC, src), "-o",
                          _obj(src, tag)])
    with ThreadPoolExecutor(jobs) as ex:
        for msg in ex.map(_run, tasks):
            if verbose and msg:
                print(msg, file=sys.stderr)
    common_objs = [_obj(s) for s in HIP_SOURCES + CPP_SOURCES]
    outs = []
    for tag in ("_f32", "_f64"):
        out = os.path.join(HERE, f"libmatgpu{tag}.so")
        tmp = out + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *common_objs,
              *[_obj(src, tag) for src in PER_TYPE]])
        os.replace(tmp, out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)
