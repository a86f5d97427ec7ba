"""The CPU oracle and the library's pure host helpers built with
AddressSanitizer and UBSan (host code only; nothing here touches a GPU) and
run over many shapes by tests/c_caller/sanitize_driver.c: out-of-bounds
accesses, leaks and undefined behaviour abort the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_oracle_and_host_helpers_under_sanitizers(tmp_path, prec):
    cc, cxx = shutil.which("gcc"), shutil.which("g++")
    if not cc or not cxx:
        pytest.skip("no host compiler")
    defs = ["-DORACLE_F64", "-DMAT_DOUBLE_PRECISION"] if prec == "f64" else []
    inc = [f"-I{os.path.join(ROOT, 'include')}"]
    objs = []
    for src, comp, extra in ((os.path.join(ROOT, "oracle", "oracle.c"), cc,
                              ["-std=c99", "-ffp-contract=off"]),
                             (os.path.join(ROOT, "libmat_amd", "csrc", "mat_host.cpp"), cxx,
                              ["-std=c++17"]),
                             (os.path.join(HERE, "c_caller", "sanitize_driver.c"), cc, ["-std=c99"])):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        r = subprocess.run([comp, *SAN, *defs, *inc, *extra, "-c", src, "-o", obj],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        objs.append(obj)
    exe = str(tmp_path / "sanitize_driver")
    r = subprocess.run([cxx, *SAN, *objs, "-lm", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # the environment is passed on as it is; a library preloaded ahead of
    # the ASan runtime is tolerated rather than removed
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "sanitize_driver: ok" in r.stdout
