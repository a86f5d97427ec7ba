"""A C program written against libmat's own header links to the GPU
library: mat.h included without MAT_IMPLEMENTATION only declares the
functions (mat.h:45-47, 359-361), so every libmat symbol the program uses
(tests/c_caller/libmat_caller.c references all that the library provides)
must resolve in libmatgpu_f32.so / libmatgpu_f64.so.  Compiles and links
only; tests/test_gpu_threads.py runs the program on the GPU."""
import os
import shutil
import subprocess

import pytest

import libmat_amd

REF_HEADER = "/root/reference/mat.h"
SRC = os.path.join(libmat_amd.ROOT, "tests", "c_caller", "libmat_caller.c")
pytestmark = pytest.mark.skipif(not libmat_amd.built(), reason="library not built")


def build_caller(out, prec, header_dir, header):
    cc = shutil.which("gcc") or shutil.which("cc")
    assert cc, "no C compiler"
    lib = f"libmatgpu_{prec}.so"
    cmd = [cc, "-O1", "-Wall", f"-DMAT_HEADER=\"{header}\"", f"-I{header_dir}", SRC,
           f"-L{libmat_amd.HERE}", f"-l:{lib}", f"-Wl,-rpath,{libmat_amd.HERE}",
           "-Wl,--no-allow-shlib-undefined", "-lpthread", "-lm", "-o", out]
    if prec == "f64":
        cmd.insert(1, "-DMAT_DOUBLE_PRECISION")
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_links_against_reference_header(tmp_path, prec):
    if not os.path.exists(REF_HEADER):
        pytest.skip("the reference header is not on this machine")
    build_caller(str(tmp_path / "caller"), prec, os.path.dirname(REF_HEADER), "mat.h")


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_links_against_own_header(tmp_path, prec):
    build_caller(str(tmp_path / "caller"), prec, os.path.join(libmat_amd.ROOT, "include"),
                 "mat_gpu.h")
