"""The drop-in entry points from several host threads at once.

libmat's scalar path is re-entrant; the GPU library gives every host thread
its own stream and workspaces, so concurrent calls neither share device
buffers nor race on the library's tables.  Checked two ways: a compiled C
caller of libmat's API (tests/c_caller/libmat_caller.c: known answers, then
threads that each solve their own systems), and Python threads (ctypes
drops the GIL during a call, so the calls overlap) whose results must equal
the oracle's on the same inputs, as in single-threaded use."""
import os
import subprocess
import threading

import numpy as np
import pytest

import libmat_amd
from _util import EPS, backward_error, pivots_match
from test_c_caller_link import build_caller

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_c_caller_runs(gpu, tmp_path, prec):
    exe = build_caller(str(tmp_path / "caller"), prec,
                       os.path.join(libmat_amd.ROOT, "include"), "mat_gpu.h")
    for threads in (1, 4):
        print(f"libmat_caller {prec} threads={threads}", flush=True)
        r = subprocess.run([exe, str(threads)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "libmat_caller: ok" in r.stdout


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_concurrent_plu_and_solve(gpu, prec):
    from oracle.oracle import Oracle
    g, o = gpu[prec], Oracle(prec)
    dt = np.float32 if prec == "f32" else np.float64
    # sizes on both sides of the look-ahead threshold (1024), so two threads
    # also share the look-ahead's second stream
    jobs = []
    for t, n in enumerate((200, 1100, 333, 1030)):
        rng = np.random.default_rng(100 + t)
        A = rng.uniform(-1, 1, (n, n)).astype(dt)
        b = rng.uniform(-1, 1, n).astype(dt)
        jobs.append((A, b))
    want = [(o.plu(A), o.solve(A, b)) for A, b in jobs]
    got = [None] * len(jobs)
    errors = []

    def work(i):
        try:
            A, b = jobs[i]
            res = []
            for _ in range(3):
                res.append((g.mat_plu(A), g.mat_solve(A, b)))
            got[i] = res
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(jobs))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=300)
    assert not errors, errors
    for i, ((A, b), ((Lo, Uo, po, so), xo)) in enumerate(zip(jobs, want)):
        n = A.shape[0]
        for (L, U, p, s), x in got[i]:
            ok, k = pivots_match(p, po, Lo, prec)
            assert ok, (n, k)
            if k is None:
                assert s == so
            assert backward_error(A, x, b) <= 8 * n * EPS[prec], n
        # repeated calls give the same bits
        (L0, U0, p0, _), x0 = got[i][0]
        for (L, U, p, _), x in got[i][1:]:
            assert np.array_equal(p, p0) and np.array_equal(U, U0) and np.array_equal(x, x0)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_c_fixture_check_on_gpu(gpu, tmp_path, prec):
    """The C fixture checker, linked to the GPU library, reproduces the
    reference's recorded outputs within the stated tolerances."""
    from test_fixture_format import GOLDEN, build_check
    exe = build_check(str(tmp_path / "fx"), prec, reference=False)
    print(f"fixture_check {prec}", flush=True)
    r = subprocess.run([exe, os.path.join(GOLDEN, f"golden_{prec}.lmfx"), "tol"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failures" in r.stdout
