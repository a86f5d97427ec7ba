"""Parity at the benchmark configurations' full sizes, and two drop-in
semantics checked on the GPU.

* Mixed-precision general solve, N = 16384, 16 right-hand sides, A = U
  diag(sigma) V^T with kappa in {1e2, 1e4, 1e6}: normwise backward error
  <= 1e-14, and forward error against the constructed solution
  x = V diag(sigma)^-1 U^T b (fp64) within FERR(kappa, n) below; the
  kappa = 1e4 solve repeated gives the same bits.
* fp32 Cholesky solve, N = 8192, 8 right-hand sides, A = G G^T / N + I:
  backward error <= n * eps32.
* Batched uneven solves, 200000 systems (n log-uniform on [3, 64], half
  general, half SPD): every system against the oracle's per-matrix
  mat_solve / mat_solve_spd (mat.h:4807-4871) -- same info, solutions
  within 1e-3 of the oracle's (relative, max norm), backward error
  <= 8 n eps32.
* Mixed-precision SPD solve, N = 32768, kappa = 1e2, 4 right-hand sides:
  backward error <= 1e-14, forward error against Q diag(lambda)^-1 Q^T b
  within FERR(kappa, n).
* mat_chol reads only the lower triangle (mat.h:6072-6082); mat_gemm with
  beta = 0 never reads C (mat.h:994-998)."""
import ctypes as C

import numpy as np
import pytest
import torch

import libmat_amd
from _util import EPS, stream

pytestmark = pytest.mark.gpu


def ptr(t):
    return C.c_void_p(t.data_ptr())


def ferr_bound(kappa, n):
    """Forward-error bound of a solve refined to normwise backward error
    beta = 1e-14 (infinity norms) on a matrix built in fp64: the computed A
    differs from the exact product by about sqrt(n) eps64 relative (a sum of
    n rounded terms), and kappa_inf <= sqrt(n) kappa_2, so
    ||x - x_true|| / ||x_true|| <~ kappa sqrt(n) (beta + sqrt(n) eps64) /
    sqrt(n); taken as kappa * sqrt(n) * 1.1e-14."""
    return kappa * np.sqrt(n) * 1.1e-14


def mixed(kind, n, nrhs, A, B):
    """A: (n, n) fp64 tensor read as a column-major buffer; B: (nrhs, n)."""
    g = libmat_amd.load("f64")
    opts = libmat_amd.IrOpts()
    g.lib.mat_gpu_ir_defaults(C.byref(opts), 1 if kind == "posv" else 0)
    st = libmat_amd.IrStatus()
    X = torch.zeros(nrhs, n, dtype=torch.float64, device="cuda")
    rc = getattr(g.lib, f"mat_gpu_{kind}_mixed")(stream(), n, nrhs, ptr(A), n, ptr(B), n, ptr(X),
                                                 n, C.byref(opts), C.byref(st))
    g.check(rc, kind)
    torch.cuda.synchronize()
    return rc, X, st


def backward_error_dev(Am, Xm, Bm):
    """Normwise ||A x - b|| / (||A|| ||x|| + ||b||) per column, infinity
    norms, fp64 on the device; Am, Xm, Bm are the logical matrices."""
    R = Am @ Xm - Bm
    nA = Am.abs().sum(dim=1).max()
    return float((R.abs().amax(dim=0) / (nA * Xm.abs().amax(dim=0) + Bm.abs().amax(dim=0))).max())


def test_mixed_lu_16384_all_kappas(gpu):
    n, nrhs = 16384, 16
    gen = torch.Generator(device="cuda").manual_seed(2026)
    U, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen))
    V, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen))
    B = torch.rand(nrhs, n, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    Bm = B.T
    for kappa in (1e2, 1e4, 1e6):
        print(f"mixed LU n={n} kappa={kappa:g}", flush=True)
        sig = torch.logspace(0, -np.log10(kappa), n, dtype=torch.float64, device="cuda")
        # the row-major tensor M = U S V^T is the column-major buffer of
        # A = M^T = V S U^T, whose solution is U S^-1 V^T b
        M = ((U * sig) @ V.T).contiguous()
        rc, X, st = mixed("gesv", n, nrhs, M, B)
        Am, Xm = M.T, X.T
        be = backward_error_dev(Am, Xm, Bm)
        Xt = U @ ((V.T @ Bm) / sig[:, None])
        fe = float((Xm - Xt).abs().max() / Xt.abs().max())
        print(f"  iters={st.iterations} fell_back={st.fell_back} berr={be:.2e} ferr={fe:.2e}",
              flush=True)
        assert rc == 0
        assert be <= 1e-14 and st.backward_error <= 1e-14
        assert fe <= ferr_bound(kappa, n)
        if kappa == 1e4:
            # the factorization is deterministic although its panels run
            # beside the trailing updates on another stream: the same
            # system solved again gives the same bits, iterations and no
            # fallback (a race between the wavefronts of one tournament
            # workgroup once changed the pivots in about half the runs)
            for rep in range(3):
                rc2, X2, st2 = mixed("gesv", n, nrhs, M, B)
                print(f"  repeat {rep}: iters={st2.iterations} fell_back={st2.fell_back}", flush=True)
                assert rc2 == 0 and st2.fell_back == 0 and st2.iterations == st.iterations
                assert torch.equal(X2, X)
                del X2
        del M, X, Xt
        torch.cuda.empty_cache()


def test_chol_8192_8rhs(gpu):
    g = gpu["f32"]
    n, nrhs = 8192, 8
    gen = torch.Generator(device="cuda").manual_seed(11)
    G = torch.randn(n, n, device="cuda", generator=gen)
    A = (G @ G.T / n + torch.eye(n, device="cuda")).contiguous()
    del G
    B = (torch.rand(nrhs, n, device="cuda", generator=gen) * 2 - 1).contiguous()
    X = torch.zeros_like(B)
    rc = g.lib.mat_gpu_posv_f32(stream(), n, nrhs, ptr(A), n, ptr(B), n, ptr(X), n)
    g.check(rc, "posv")
    torch.cuda.synchronize()
    assert rc == 0
    be = backward_error_dev(A.double(), X.double().T, B.double().T)
    print(f"cholesky n={n} berr={be:.2e}", flush=True)
    assert be <= n * EPS["f32"]


def make_batch_fast(count, seed, lo=3, hi=64):
    """The batched configuration's systems, generated per size at once."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(lo), np.log(hi + 1), count)).astype(np.int32).clip(lo, hi)
    kinds = (np.arange(count) % 2).astype(np.int8)
    aoff = np.zeros(count, np.int64)
    voff = np.zeros(count, np.int64)
    aoff[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
    voff[1:] = np.cumsum(ns.astype(np.int64))[:-1]
    A = np.empty(int(aoff[-1] + ns[-1] ** 2), np.float32)
    b = rng.uniform(-1, 1, int(voff[-1] + ns[-1])).astype(np.float32)
    for n in range(lo, hi + 1):
        for k in (0, 1):
            idx = np.nonzero((ns == n) & (kinds == k))[0]
            if idx.size == 0:
                continue
            if k == 0:
                M = rng.uniform(-1, 1, (idx.size, n, n)) + n * np.eye(n)
            else:
                G = rng.standard_normal((idx.size, n, n))
                M = G @ np.transpose(G, (0, 2, 1)) + np.eye(n)
            Mf = np.transpose(M, (0, 2, 1)).reshape(idx.size, n * n).astype(np.float32)
            A[(aoff[idx][:, None] + np.arange(n * n)[None, :]).ravel()] = Mf.ravel()
    return ns, kinds, A, b, aoff, voff


def test_batched_200k_against_oracle(gpu):
    from oracle.oracle import Oracle
    g = gpu["f32"]
    count = 200000
    ns, kinds, A, b, aoff, voff = make_batch_fast(count, seed=5)
    print(f"batched count={count}", flush=True)
    xo, io = Oracle("f32").batch_solve(ns, kinds, A, aoff, b, voff)
    d = {k: torch.from_numpy(v).cuda() for k, v in (("A", A), ("b", b), ("aoff", aoff),
                                                      ("voff", voff))}
    x = torch.zeros(b.size, dtype=torch.float32, device="cuda")
    info = torch.full((count,), 7, dtype=torch.int32, device="cuda")
    s = stream()
    plan = C.c_void_p()
    g.check(g.lib.mat_gpu_batch_create_f32(s, count, ns.ctypes.data_as(C.c_void_p),
                                           kinds.ctypes.data_as(C.c_void_p), ptr(d["A"]),
                                           ptr(d["aoff"]), ptr(d["b"]), ptr(x), ptr(d["voff"]),
                                           ptr(info), 0, C.byref(plan)), "plan")
    try:
        g.check(g.lib.mat_gpu_batch_run(s, plan), "run")
        torch.cuda.synchronize()
    finally:
        g.lib.mat_gpu_batch_destroy(plan)
    xg, ig = x.cpu().numpy(), info.cpu().numpy()
    assert np.array_equal(ig, io)
    worst_f = worst_b = 0.0
    for n in range(3, 65):
        idx = np.nonzero(ns == n)[0]
        if idx.size == 0:
            continue
        M = A[(aoff[idx][:, None] + np.arange(n * n)[None, :])].reshape(-1, n, n)
        M = np.transpose(M, (0, 2, 1)).astype(np.float64)  # column-major -> (row, col)
        vi = (voff[idx][:, None] + np.arange(n)[None, :])
        X, Xo, Bv = xg[vi].astype(np.float64), xo[vi].astype(np.float64), b[vi].astype(np.float64)
        fe = np.abs(X - Xo).max(axis=1) / np.abs(Xo).max(axis=1)
        R = np.einsum("kij,kj->ki", M, X) - Bv
        be = np.abs(R).max(axis=1) / (np.abs(M).sum(axis=2).max(axis=1) * np.abs(X).max(axis=1)
                                      + np.abs(Bv).max(axis=1))
        assert fe.max() <= 1e-3, (n, fe.max())
        assert be.max() <= 8 * n * EPS["f32"], (n, be.max())
        worst_f, worst_b = max(worst_f, fe.max()), max(worst_b, be.max())
    print(f"  worst forward vs oracle {worst_f:.2e}, worst backward {worst_b:.2e}", flush=True)


def test_mixed_spd_32768_kappa_1e2(gpu):
    n, nrhs, kappa = 32768, 4, 1e2
    gen = torch.Generator(device="cuda").manual_seed(2027)
    Q = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen)
    for _ in range(2):  # CholeskyQR2: Q R^-1 with R = chol(Q^T Q), twice
        R = torch.linalg.cholesky(Q.T @ Q, upper=True)
        Q = torch.linalg.solve_triangular(R, Q, upper=True, left=False)
        del R
    lam = torch.exp(torch.rand(n, dtype=torch.float64, device="cuda", generator=gen)
                    * np.log(kappa))
    A = (Q * lam) @ Q.T
    A = ((A + A.T) * 0.5).contiguous()
    B = torch.randn(nrhs, n, dtype=torch.float64, device="cuda", generator=gen)
    print(f"mixed SPD n={n} kappa={kappa:g}", flush=True)
    rc, X, st = mixed("posv", n, nrhs, A, B)
    be = backward_error_dev(A, X.T, B.T)
    Xt = Q @ ((Q.T @ B.T) / lam[:, None])
    fe = float((X.T - Xt).abs().max() / Xt.abs().max())
    print(f"  iters={st.iterations} fell_back={st.fell_back} precision={st.precision} "
          f"berr={be:.2e} ferr={fe:.2e}", flush=True)
    assert rc == 0
    assert be <= 1e-14 and st.backward_error <= 1e-14
    assert fe <= ferr_bound(kappa, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_chol_reads_only_the_lower_triangle(gpu, prec):
    from oracle.oracle import Oracle
    g, o = gpu[prec], Oracle(prec)
    dt = np.float32 if prec == "f32" else np.float64
    rng = np.random.default_rng(8)
    for n in (5, 64, 300, 1100):
        G = rng.standard_normal((n, n))
        S = (G @ G.T / n + np.eye(n)).astype(dt)
        junk = S.copy()
        iu = np.triu_indices(n, 1)
        junk[iu] = np.nan if n % 2 else 1e30
        L0, r0 = g.mat_chol(S)
        L1, r1 = g.mat_chol(junk)
        assert r0 == r1 == 0
        assert np.array_equal(L0, L1), n
        Lo, ro = o.chol(junk)  # the oracle reads the lower triangle as well
        assert ro == 0
        assert np.abs(L1 - Lo).max() <= 8 * n * EPS[prec] * np.abs(Lo).max(), n
        b = rng.uniform(-1, 1, n).astype(dt)
        x0, s0 = g.mat_solve_spd(S, b)
        x1, s1 = g.mat_solve_spd(junk, b)
        assert s0 == s1 == 0 and np.array_equal(x0, x1), n


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_gemm_beta_zero_never_reads_c(gpu, prec):
    g = gpu[prec]
    dt = np.float32 if prec == "f32" else np.float64
    rng = np.random.default_rng(9)
    for m, k, n in ((3, 4, 5), (130, 70, 257), (1024, 512, 640)):
        A = rng.uniform(-1, 1, (m, k)).astype(dt)
        B = rng.uniform(-1, 1, (k, n)).astype(dt)
        Cn = np.full((m, n), np.nan, dt)
        out = g.mat_gemm(1.0, A, B, 0.0, Cn)
        assert not np.isnan(out).any(), (m, k, n)
        assert np.array_equal(out, g.mat_mul(A, B)), (m, k, n)
