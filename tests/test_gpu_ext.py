"""Device-resident extension entry points (include/mat_gpu.h part 2) at
sizes past the drop-in tests: MFMA GEMM shapes and transposes, the
multi-workgroup pivoting path, multi-RHS solves, and full-size properties
(sampled fp64 dot products for the 8192^3 GEMM)."""
import ctypes as C

import numpy as np
import pytest
import torch

from _util import DT, EPS, apply_row_perm, backward_error, from_dev, pivots_match, rel_fro, \
    stream, to_dev
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
PRECS = ["f32", "f64"]


def ptr(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("m,n,k,ta,tb", [(1, 1, 1, 0, 0), (129, 70, 33, 0, 0), (64, 256, 200, 1, 0),
                                         (300, 17, 128, 0, 1), (5, 400, 9, 1, 1),
                                         (513, 257, 1000, 0, 0)])
def test_gemm_shapes(gpu, prec, m, n, k, ta, tb):
    g = gpu[prec]
    rng = np.random.default_rng(m * 7 + n)
    A = rng.uniform(-1, 1, (k, m) if ta else (m, k))
    B = rng.uniform(-1, 1, (n, k) if tb else (k, n))
    Cm = rng.uniform(-1, 1, (m, n))
    dA, dB, dC = to_dev(A, DT[prec]), to_dev(B, DT[prec]), to_dev(Cm, DT[prec])
    fn = getattr(g.lib, f"mat_gpu_gemm_{prec}")
    rc = fn(stream(), ta, tb, m, n, k, 1.25, ptr(dA), A.shape[0], ptr(dB), B.shape[0], -0.5,
            ptr(dC), m)
    g.check(rc, "gemm")
    torch.cuda.synchronize()
    ref = 1.25 * (A.T if ta else A) @ (B.T if tb else B) - 0.5 * Cm
    assert rel_fro(from_dev(dC, m, n), ref) <= 2 * k * EPS[prec]


def test_gemm_8192_sampled():
    """M=N=K=8192: 256 entries recomputed as fp64 dot products on the host."""
    import libmat_amd
    g = libmat_amd.load("f32")
    n = 8192
    gen = torch.Generator(device="cuda").manual_seed(1234)
    dA = (torch.rand(n * n, device="cuda", generator=gen) * 2 - 1)
    dB = (torch.rand(n * n, device="cuda", generator=gen) * 2 - 1)
    dC = torch.empty(n * n, device="cuda")
    g.check(g.lib.mat_gpu_gemm_f32(stream(), 0, 0, n, n, n, 1.0, ptr(dA), n, ptr(dB), n, 0.0,
                                   ptr(dC), n), "gemm")
    torch.cuda.synchronize()
    rng = np.random.default_rng(9)
    ii, jj = rng.integers(0, n, 256), rng.integers(0, n, 256)
    A = dA.view(n, n)  # [col][row] since storage is column-major
    B = dB.view(n, n)
    rows = A[:, torch.as_tensor(ii, device="cuda")].T.double().cpu().numpy()  # A[i, :]
    cols = B[torch.as_tensor(jj, device="cuda"), :].double().cpu().numpy()  # B[:, j]
    ref = np.einsum("ik,ik->i", rows, cols)
    got = dC.view(n, n)[torch.as_tensor(jj, device="cuda"), torch.as_tensor(ii, device="cuda")]
    got = got.double().cpu().numpy()
    scale = np.sqrt(np.einsum("ik,ik->i", rows ** 2, cols ** 2))
    assert np.max(np.abs(got - ref) / scale) <= 2 * np.sqrt(n) * EPS["f32"] * 8


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n,leaf", [(300, 32), (1000, 24), (1000, 32), (1100, 32), (2500, 32),
                                    (100, 64), (300, 64), (1100, 64)])
def test_getrf_vs_oracle(gpu, prec, n, leaf):
    """Multi-workgroup pivoting (m > 256 rows) against the oracle's PLU; from
    n = 1024 (leaf 32) the look-ahead driver with its second stream runs."""
    g = gpu[prec]
    rng = np.random.default_rng(n + leaf)
    A = rng.uniform(-1, 1, (n, n)).astype(DT[prec])
    dA = to_dev(A, DT[prec])
    perm = torch.zeros(n, dtype=torch.int32, device="cuda")
    sw = C.c_int(0)
    rc = getattr(g.lib, f"mat_gpu_getrf_{prec}")(stream(), n, ptr(dA), n, ptr(perm), leaf,
                                                 C.byref(sw))
    g.check(rc, "getrf")
    torch.cuda.synchronize()
    LU = from_dev(dA, n, n).astype(np.float64)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    p = perm.cpu().numpy()
    assert sorted(p.tolist()) == list(range(n))
    assert rel_fro(L @ U, apply_row_perm(A, p)) <= 8 * n * EPS[prec]
    if n <= 1100:
        Lr, Ur, pr, sr = Oracle(prec).plu(A)
        ok, k = pivots_match(p, pr, Lr, prec)
        assert ok, f"pivot mismatch at step {k}"
        if k is None:
            assert sw.value == sr
    # solve with the packed factor, 5 right-hand sides
    B = rng.uniform(-1, 1, (n, 5))
    dB, dX = to_dev(B, DT[prec]), torch.zeros(n * 5, dtype=dA.dtype, device="cuda")
    g.check(getattr(g.lib, f"mat_gpu_getrs_{prec}")(stream(), n, 5, ptr(dA), n, ptr(perm), ptr(dB),
                                                    n, ptr(dX), n), "getrs")
    torch.cuda.synchronize()
    assert backward_error(A, from_dev(dX, n, 5), B) <= 8 * n * EPS[prec]


def _perm_parity(p):
    """(-1)^(number of transpositions) of the permutation p."""
    p = list(p)
    seen, par = [False] * len(p), 0
    for i in range(len(p)):
        if not seen[i]:
            j, c = i, 0
            while not seen[j]:
                seen[j] = True
                j = p[j]
                c += 1
            par ^= (c - 1) & 1
    return par


@pytest.mark.parametrize("n", [32, 100, 300, 1100, 2050, 3000, 9000, 16500])
def test_getrf_tournament(gpu, n):
    """Tournament pivoting (fp32), one workgroup per 256 rows and eight
    workgroups per node of the reduction tree: one workgroup (n <= 256), one
    level above the leaves (n = 300: 2 workgroups; 1100: 5), two levels
    (3000: 12 workgroups, look-ahead driver; 9000: 36) and three (16500: 65
    workgroups, a ragged last node); 2050 has rows that are not 16-byte
    aligned, so its leaves load and store their rows element by element
    instead of in 16-byte vectors.  Checked by what the mathematics fixes, since
    the pivots are not libmat's: perm is a permutation whose parity the
    swap count gives, PA = LU to fp32 rounding (fp64 on the GPU), the
    multipliers stay small (the tournament's growth is close to partial
    pivoting's), a solve with the factor, and bit-identical results on a
    second run (every workgroup plays the last round itself; they must
    agree)."""
    g = gpu["f32"]
    gen = torch.Generator(device="cuda").manual_seed(n + 7)
    A0 = torch.rand(n * n, device="cuda", dtype=torch.float32, generator=gen) * 2 - 1
    outs = []
    for _ in range(2):
        A = A0.clone()
        perm = torch.zeros(n, dtype=torch.int32, device="cuda")
        sw = C.c_int(0)
        g.check(g.lib.mat_gpu_getrf_tournament_f32(stream(), n, ptr(A), n, ptr(perm), C.byref(sw)),
                "getrf_tournament")
        torch.cuda.synchronize()
        outs.append((A, perm, sw.value))
    (A, perm, sw), (A2, perm2, sw2) = outs
    assert torch.equal(A, A2) and torch.equal(perm, perm2) and sw == sw2
    del A2
    p = perm.long()
    assert torch.equal(torch.sort(p).values, torch.arange(n, device="cuda"))
    assert _perm_parity(p.cpu().tolist()) == (sw & 1)
    LU = A.view(n, n).T.double()  # column-major buffer -> logical matrix
    L = torch.tril(LU, -1) + torch.eye(n, device="cuda", dtype=torch.float64)
    U = torch.triu(LU)
    del LU
    lmax = float(L.abs().max())
    PA = A0.view(n, n).T.double()[p]
    err = float(torch.linalg.matrix_norm(L @ U - PA) / torch.linalg.matrix_norm(PA))
    print(f"n={n} max|L|={lmax:.2f} swaps={sw} rel err {err:.2e}")
    del L, U, PA
    assert lmax <= 8.0
    assert err <= 8 * n * EPS["f32"], err
    B = torch.rand(3 * n, device="cuda", dtype=torch.float32, generator=gen) * 2 - 1
    X = torch.zeros_like(B)
    g.check(g.lib.mat_gpu_getrs_f32(stream(), n, 3, ptr(A), n, ptr(perm), ptr(B), n, ptr(X), n),
            "getrs")
    torch.cuda.synchronize()
    Am = A0.view(n, n).T.double()
    Xm, Bm = X.view(3, n).T.double(), B.view(3, n).T.double()
    be = float(((Am @ Xm - Bm).abs().amax(0) /
                (Am.abs().sum(1).max() * Xm.abs().amax(0) + Bm.abs().amax(0))).max())
    assert be <= 8 * n * EPS["f32"], be


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n", [100, 700, 2049])
def test_gesv_posv(gpu, prec, n):
    g = gpu[prec]
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n))
    G = rng.standard_normal((n, n))
    S = G @ G.T / n + np.eye(n)
    B = rng.uniform(-1, 1, (n, 8))
    for fn, M in (("gesv", A), ("posv", S)):
        dA, dB = to_dev(M, DT[prec]), to_dev(B, DT[prec])
        dX = torch.zeros(n * 8, dtype=dA.dtype, device="cuda")
        rc = getattr(g.lib, f"mat_gpu_{fn}_{prec}")(stream(), n, 8, ptr(dA), n, ptr(dB), n,
                                                    ptr(dX), n)
        assert g.check(rc, fn) == 0
        torch.cuda.synchronize()
        assert backward_error(M.astype(DT[prec]), from_dev(dX, n, 8), B) <= 8 * n * EPS[prec], fn
        # A must be untouched
        assert np.array_equal(from_dev(dA, n, n), M.astype(DT[prec]))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n", [300, 1300])
def test_potrf_not_spd(gpu, prec, n):
    """n = 1300: the blocked look-ahead driver (blocks of 256, the failing
    column in the fifth block, after the side stream's updates began)."""
    g = gpu[prec]
    rng = np.random.default_rng(1)
    G = rng.standard_normal((n, n))
    S = G @ G.T / n + np.eye(n)
    bad = 2 * n // 3
    S[bad, bad] = -5.0  # fails at (or before) that column
    dA = to_dev(S, DT[prec])
    assert getattr(g.lib, f"mat_gpu_potrf_{prec}")(stream(), n, ptr(dA), n) == -1
    dA = to_dev(G @ G.T / n + np.eye(n), DT[prec])
    assert getattr(g.lib, f"mat_gpu_potrf_{prec}")(stream(), n, ptr(dA), n) == 0
    L = from_dev(dA, n, n)
    assert np.all(np.triu(L, 1) == 0)
    assert rel_fro(L.astype(np.float64) @ L.T, G @ G.T / n + np.eye(n)) <= 8 * n * EPS[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("leaf", [32, 64])
def test_getrf_tall_panel(gpu, prec, leaf):
    """n = 16500 > 16384: the first leaves hold two rows per thread (64
    workgroups of 256 threads cover 16384 rows with one); with 64-column
    leaves f64 blocks taller than that use 32-column leaves.  PA = LU
    checked in fp64 on the GPU."""
    g = gpu[prec]
    n = 16500
    dt = torch.float32 if prec == "f32" else torch.float64
    gen = torch.Generator(device="cuda").manual_seed(77)
    A = (torch.rand(n * n, device="cuda", dtype=dt, generator=gen) * 2 - 1)
    A0 = A.clone()
    perm = torch.zeros(n, dtype=torch.int32, device="cuda")
    sw = C.c_int(0)
    g.check(getattr(g.lib, f"mat_gpu_getrf_{prec}")(stream(), n, ptr(A), n, ptr(perm), leaf,
                                                     C.byref(sw)), "getrf")
    torch.cuda.synchronize()
    LU = A.view(n, n).T.double()  # column-major buffer -> logical matrix
    p = perm.long()
    assert torch.equal(torch.sort(p).values, torch.arange(n, device="cuda"))
    L = torch.tril(LU, -1) + torch.eye(n, device="cuda", dtype=torch.float64)
    U = torch.triu(LU)
    del LU
    PA = A0.view(n, n).T.double()[p]
    err = float(torch.linalg.matrix_norm(L @ U - PA) / torch.linalg.matrix_norm(PA))
    assert err <= 8 * n * EPS[prec], err


@pytest.mark.parametrize("prec", PRECS)
def test_geqrf_device_matches_dropin(gpu, prec):
    """mat_gpu_geqrf (device pointers) runs mat_qr's kernels: R and Q bit for
    bit those of the drop-in mat_qr, and with q = NULL the same R (the
    mat_qr_r form); m > n with a ragged last block of reflectors."""
    g = gpu[prec]
    m, n = 700, 333
    A = np.random.default_rng(4).uniform(-1, 1, (m, n))
    Qd, Rd = g.mat_qr(A.astype(DT[prec]))
    fn = getattr(g.lib, f"mat_gpu_geqrf_{prec}")
    dA = to_dev(A, DT[prec])
    dQ = to_dev(np.zeros((m, m)), DT[prec])
    assert fn(stream(), m, n, ptr(dA), m, ptr(dQ)) == 0
    assert np.array_equal(from_dev(dA, m, n), Rd)
    assert np.array_equal(from_dev(dQ, m, m), Qd)
    dA = to_dev(A, DT[prec])
    assert fn(stream(), m, n, ptr(dA), m, None) == 0
    assert np.array_equal(from_dev(dA, m, n), Rd)


@pytest.mark.parametrize("prec", PRECS)
def test_geqrf_padded_leading_dimension(gpu, prec):
    """A padded leading dimension (lda = m + 7, which the header allows):
    R (the top m rows) and Q bit for bit the lda = m result, with and
    without q, and the padding rows untouched."""
    g = gpu[prec]
    m, n, pad = 700, 333, 7
    A = np.random.default_rng(5).uniform(-1, 1, (m, n))
    fn = getattr(g.lib, f"mat_gpu_geqrf_{prec}")
    dA = to_dev(A, DT[prec])
    dQ = to_dev(np.zeros((m, m)), DT[prec])
    assert fn(stream(), m, n, ptr(dA), m, ptr(dQ)) == 0
    R0, Q0 = from_dev(dA, m, n), from_dev(dQ, m, m)
    for with_q in (True, False):
        Ap = np.full((m + pad, n), 12345.0)
        Ap[:m] = A
        dAp = to_dev(Ap, DT[prec])
        dQp = to_dev(np.zeros((m, m)), DT[prec])
        assert fn(stream(), m, n, ptr(dAp), m + pad, ptr(dQp) if with_q else None) == 0
        got = from_dev(dAp, m + pad, n)
        assert np.array_equal(got[:m], R0), with_q
        assert np.all(got[m:] == 12345.0), with_q
        if with_q:
            assert np.array_equal(from_dev(dQp, m, m), Q0)


def test_geqrf_full_size_properties(gpu):
    """QR at the benchmark's size (fp32 4096 x 4096, through the device
    entry point): Q^T Q = I and A = Q R checked in fp64 on the GPU, the
    residue below R's diagonal within the reference test's m * eps bound,
    and the R-only call (q = NULL) giving the same R bit for bit."""
    g = gpu["f32"]
    n = 4096
    gen = torch.Generator(device="cuda").manual_seed(17)
    A0 = (torch.rand(n * n, device="cuda", generator=gen) * 2 - 1).contiguous()
    A = A0.clone()
    Q = torch.empty_like(A0)
    assert g.lib.mat_gpu_geqrf_f32(stream(), n, n, ptr(A), n, ptr(Q)) == 0
    M = A0.view(n, n).double().T
    Qm = Q.view(n, n).double().T
    Rm = A.view(n, n).double().T
    eye = torch.eye(n, dtype=torch.float64, device="cuda")
    assert float((Qm.T @ Qm - eye).abs().max()) <= 20 * n * EPS["f32"]
    assert float((Qm @ torch.triu(Rm) - M).norm() / M.norm()) <= 20 * n * EPS["f32"]
    assert float(torch.tril(Rm, -1).abs().max()) <= n * 1e-6
    A2 = A0.clone()
    assert g.lib.mat_gpu_geqrf_f32(stream(), n, n, ptr(A2), n, None) == 0
    torch.cuda.synchronize()
    assert torch.equal(A2, A)
