"""mat_gpu_batch_solve_noalloc_f32: the batched solve with its scratch in a
fixed per-GPU array (no allocation) gives the same bits as
mat_gpu_batch_solve_f32 on the same inputs, on two calls in a row (the
second smaller, so stale scratch from the first would show), including a
matrix that is not positive definite (info -1, x untouched); batches above
the array's limit are refused with a message."""
import ctypes as C

import numpy as np
import pytest
import torch

from test_gpu_batched import make_batch, ptr, upload

pytestmark = pytest.mark.gpu


def solve(g, fn, ns, kinds, d):
    x = torch.full_like(d["x"], 3.5)  # sentinel: stays where nothing is written
    info = torch.full_like(d["info"], 7)
    s = torch.cuda.current_stream().cuda_stream
    rc = getattr(g.lib, fn)(s, ns.size, ns.ctypes.data_as(C.c_void_p),
                            kinds.ctypes.data_as(C.c_void_p), ptr(d["A"]), ptr(d["aoff"]),
                            ptr(d["b"]), ptr(x), ptr(d["voff"]), ptr(info))
    g.check(rc, fn)
    assert rc == 0
    torch.cuda.synchronize()
    return x.cpu().numpy(), info.cpu().numpy()


def test_noalloc_matches_batch_solve(gpu):
    g = gpu["f32"]
    for count, seed in ((3000, 41), (700, 42)):
        print(f"noalloc batch count={count} seed={seed}", flush=True)
        ns, kinds, mats, bs, A, b, aoff, voff = make_batch(count, seed)
        # one SPD member that is not positive definite: -1, x untouched
        j = int(np.nonzero(kinds == 1)[0][3])
        n = int(ns[j])
        A[aoff[j]:aoff[j] + n * n] *= -1.0
        d = upload(ns, A, b, aoff, voff)
        x0, i0 = solve(g, "mat_gpu_batch_solve_f32", ns, kinds, d)
        x1, i1 = solve(g, "mat_gpu_batch_solve_noalloc_f32", ns, kinds, d)
        assert np.array_equal(i0, i1)
        assert i1[j] == -1 and np.count_nonzero(i1) == 1
        assert np.all(x1[voff[j]:voff[j] + n] == 3.5)
        assert np.array_equal(x0, x1)
        # and the solutions solve their systems
        for i in range(0, count, 97):
            if i == j:
                continue
            M = mats[i].astype(np.float64)
            xi = x1[voff[i]:voff[i] + ns[i]].astype(np.float64)
            r = np.abs(M @ xi - bs[i]).max()
            assert r <= 1e-4 * (np.abs(M).sum(1).max() * np.abs(xi).max() + np.abs(bs[i]).max())


def test_noalloc_rejects_oversized_batch(gpu):
    g = gpu["f32"]
    count = (1 << 18) + 1
    ns = np.full(count, 3, np.int32)
    kinds = np.zeros(count, np.int8)
    rc = g.lib.mat_gpu_batch_solve_noalloc_f32(None, count, ns.ctypes.data_as(C.c_void_p),
                                               kinds.ctypes.data_as(C.c_void_p), None, None,
                                               None, None, None, None)
    assert rc < 0 and "at most 262144" in g.last_error()
