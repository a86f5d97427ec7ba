"""The LMFX fixture format (tests/golden/lmfx.py, lmfx.h): Python writes
and reads it losslessly, and a C program reads the committed fixtures and
reproduces them with libmat's own API -- bit for bit when that API is the
reference itself (compiled from its header where it is present), and
within the GPU library's tolerances on the GPU (tests/test_gpu_threads.py)."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import libmat_amd

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)
import lmfx  # noqa: E402

CHECK = os.path.join(HERE, "c_caller", "fixture_check.c")
REF_DIR = "/root/reference"


def test_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    entries = {
        "m32": rng.standard_normal((3, 5)).astype(np.float32),
        "m64": rng.standard_normal((4, 1)),          # a 4 x 1 matrix stays 2-D
        "v32": rng.standard_normal(7).astype(np.float32),  # a Vec comes back 1-D
        "perm": np.array([2, 0, 1], np.int64),
        "swaps": np.int64(-3),
        "det": np.float64(-0.0),
        "nan": np.array([np.nan, np.inf, -np.inf], np.float64),
        "x" * 200: np.zeros((1, 1), np.float32),
    }
    path = str(tmp_path / "t.lmfx")
    lmfx.write(path, entries)
    back = lmfx.read(path)
    assert set(back) == set(entries)
    for k, v in entries.items():
        v = np.asarray(v)
        assert back[k].dtype == v.dtype and back[k].shape == v.shape, k
        assert np.array_equal(back[k], v, equal_nan=True), k
    assert np.signbit(back["det"])
    # payloads are 8-byte aligned within the file, column-major like Mat
    raw = open(path, "rb").read()
    assert raw[:8] == b"LMATFX01" and len(raw) % 8 == 0
    col_major = np.asfortranarray(entries["m32"]).ravel(order="F").tobytes()
    at = raw.find(col_major)
    assert at > 0 and at % 8 == 0


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_committed_fixtures_load(prec):
    g = lmfx.read(os.path.join(GOLDEN, f"golden_{prec}.lmfx"))
    assert g["plu_24_A"].shape == (24, 24) and g["plu_24_p"].dtype == np.int64
    assert sorted(g["plu_24_p"].tolist()) == list(range(24))
    assert g["notpd2_rc"] == -1 and g["chol_31_rc"] == 0


def build_check(out, prec, reference):
    cc = shutil.which("gcc") or shutil.which("cc")
    cmd = [cc, "-O2", "-w", f"-I{GOLDEN}", CHECK, "-lm", "-o", out]
    if reference:
        cmd[1:1] = ["-DCHECK_REFERENCE", f"-I{REF_DIR}"]
    else:
        lib = f"libmatgpu_{prec}.so"
        cmd[1:1] = ['-DMAT_HEADER="mat_gpu.h"', f"-I{os.path.join(libmat_amd.ROOT, 'include')}"]
        cmd += [f"-L{libmat_amd.HERE}", f"-l:{lib}", f"-Wl,-rpath,{libmat_amd.HERE}"]
    if prec == "f64":
        cmd.insert(1, "-DMAT_DOUBLE_PRECISION")
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_c_reads_fixtures_bit_exact_on_reference(tmp_path, prec):
    if not os.path.exists(os.path.join(REF_DIR, "mat.h")):
        pytest.skip("the reference header is not on this machine")
    exe = build_check(str(tmp_path / "fx"), prec, reference=True)
    r = subprocess.run([exe, os.path.join(GOLDEN, f"golden_{prec}.lmfx"), "exact"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failures" in r.stdout


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.skipif(not libmat_amd.built(), reason="library not built")
def test_c_checker_links_against_gpu_library(tmp_path, prec):
    build_check(str(tmp_path / "fx"), prec, reference=False)
