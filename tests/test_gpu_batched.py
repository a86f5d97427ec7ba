"""Batched uneven small solves: one persistent size-binned launch (and its
hipGraph replay) against per-matrix oracle solves (mat_solve /
mat_solve_spd semantics, including -1 with x untouched for a matrix that
is not positive definite)."""
import ctypes as C

import numpy as np
import pytest
import torch

import libmat_amd
from _util import stream
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def ptr(t):
    return C.c_void_p(t.data_ptr())


def make_batch(count, seed, lo=3, hi=64):
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(lo), np.log(hi + 1), count)).astype(np.int32).clip(lo, hi)
    kinds = (np.arange(count) % 2).astype(np.int8)
    mats, bs = [], []
    for n, k in zip(ns, kinds):
        if k == 0:
            M = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
        else:
            G = rng.standard_normal((n, n))
            M = G @ G.T + np.eye(n)
        mats.append(M.astype(np.float32))
        bs.append(rng.uniform(-1, 1, n).astype(np.float32))
    aoff = np.zeros(count, np.int64)
    voff = np.zeros(count, np.int64)
    aoff[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
    voff[1:] = np.cumsum(ns.astype(np.int64))[:-1]
    A = np.concatenate([m.ravel(order="F") for m in mats])
    b = np.concatenate(bs)
    return ns, kinds, mats, bs, A, b, aoff, voff


def upload(ns, A, b, aoff, voff):
    d = {"A": torch.from_numpy(A).cuda(), "b": torch.from_numpy(b).cuda(),
         "x": torch.zeros(b.size, dtype=torch.float32, device="cuda"),
         "aoff": torch.from_numpy(aoff).cuda(), "voff": torch.from_numpy(voff).cuda(),
         "info": torch.full((ns.size,), 7, dtype=torch.int32, device="cuda")}
    return d


def check(ns, kinds, mats, bs, x, info, voff):
    o = Oracle("f32")
    for i, (n, k) in enumerate(zip(ns, kinds)):
        xi = x[voff[i]:voff[i] + n]
        ref = o.solve(mats[i], bs[i]) if k == 0 else o.solve_spd(mats[i], bs[i])[0]
        assert info[i] == 0, i
        err = np.abs(xi - ref).max() / max(1e-30, np.abs(ref).max())
        assert err <= 1e-4, (i, n, k, err)


def test_batch_single_launch():
    g = libmat_amd.load("f32")
    ns, kinds, mats, bs, A, b, aoff, voff = make_batch(3000, seed=1)
    d = upload(ns, A, b, aoff, voff)
    rc = g.lib.mat_gpu_batch_solve_f32(stream(), ns.size, ns.ctypes.data_as(C.c_void_p),
                                       kinds.ctypes.data_as(C.c_void_p), ptr(d["A"]), ptr(d["aoff"]),
                                       ptr(d["b"]), ptr(d["x"]), ptr(d["voff"]), ptr(d["info"]))
    g.check(rc, "batch")
    torch.cuda.synchronize()
    check(ns, kinds, mats, bs, d["x"].cpu().numpy(), d["info"].cpu().numpy(), voff)


def test_batch_graph_replay_and_not_spd():
    g = libmat_amd.load("f32")
    ns, kinds, mats, bs, A, b, aoff, voff = make_batch(1200, seed=2)
    # matrix 1 is SPD-kind: make it indefinite
    n1 = ns[1]
    bad = mats[1].copy()
    bad[0, 0] = -1.0
    A[aoff[1]:aoff[1] + n1 * n1] = bad.ravel(order="F")
    d = upload(ns, A, b, aoff, voff)
    plan = C.c_void_p()
    rc = g.lib.mat_gpu_batch_create_f32(stream(), ns.size, ns.ctypes.data_as(C.c_void_p),
                                        kinds.ctypes.data_as(C.c_void_p), ptr(d["A"]),
                                        ptr(d["aoff"]), ptr(d["b"]), ptr(d["x"]), ptr(d["voff"]),
                                        ptr(d["info"]), 1, C.byref(plan))
    g.check(rc, "batch_create")
    try:
        for rep in range(3):
            d["x"].zero_()
            torch.cuda.synchronize()
            g.check(g.lib.mat_gpu_batch_run(stream(), plan), "batch_run")
            torch.cuda.synchronize()
            x = d["x"].cpu().numpy()
            info = d["info"].cpu().numpy()
            assert info[1] == -1 and np.all(x[voff[1]:voff[1] + n1] == 0)
            info[1] = 0
            mats_chk = list(mats)
            keep = [i for i in range(ns.size) if i != 1]
            check(ns[keep], kinds[keep], [mats_chk[i] for i in keep], [bs[i] for i in keep], x,
                  info[keep], voff[keep])
            # new right-hand sides behind the same pointers between replays
            b2 = (b * (rep + 2)).astype(np.float32)
            d["b"].copy_(torch.from_numpy(b2))
            bs = [bb * (rep + 2) / (rep + 1) for bb in bs]
        wg, om, ms = C.c_int(), C.c_int(), C.c_double()
        g.check(g.lib.mat_gpu_batch_tuning(plan, C.byref(wg), C.byref(om), C.byref(ms)), "tuning")
        assert wg.value in (1, 2, 4, 8) and ms.value > 0
    finally:
        g.lib.mat_gpu_batch_destroy(plan)


@pytest.mark.parametrize("lo,hi", [(3, 64), (3, 9), (33, 64)])
def test_tuned_plan_matches_untuned(lo, hi):
    """The autotuner's choices (bins in use, wavefronts per workgroup per
    class, workgroups per CU, class order) change where a matrix is solved,
    not how: a tuned plan gives the untuned plan's bits, and both the
    oracle's answers."""
    g = libmat_amd.load("f32")
    ns, kinds, mats, bs, A, b, aoff, voff = make_batch(2000, seed=7 + lo, lo=lo, hi=hi)
    out = []
    for tune in (0, 1):
        d = upload(ns, A, b, aoff, voff)
        plan = C.c_void_p()
        g.check(g.lib.mat_gpu_batch_create_f32(stream(), ns.size, ns.ctypes.data_as(C.c_void_p),
                                               kinds.ctypes.data_as(C.c_void_p), ptr(d["A"]),
                                               ptr(d["aoff"]), ptr(d["b"]), ptr(d["x"]),
                                               ptr(d["voff"]), ptr(d["info"]), tune,
                                               C.byref(plan)), "batch_create")
        try:
            g.check(g.lib.mat_gpu_batch_run(stream(), plan), "batch_run")
            torch.cuda.synchronize()
            mask, untuned = C.c_int(), C.c_double()
            waves = (C.c_int32 * 8)()
            g.check(g.lib.mat_gpu_batch_tuning_shape(plan, C.byref(mask), waves,
                                                     C.byref(untuned)), "shape")
            assert mask.value & 8 and all(w in (2, 4, 8) for w in waves)
            if tune:
                assert untuned.value > 0
                print(f"orders {lo}..{hi}: bins {mask.value:04b} waves {list(waves)} "
                      f"untuned {untuned.value:.3f} ms", flush=True)
        finally:
            g.lib.mat_gpu_batch_destroy(plan)
        out.append((d["x"].cpu().numpy(), d["info"].cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    check(ns, kinds, mats, bs, out[1][0], out[1][1], voff)
