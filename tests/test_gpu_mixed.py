"""Mixed-precision solvers: BF16-MFMA LU and FP8-MFMA Cholesky with fp64
iterative refinement.  Each case states backward error, forward error
against an fp64 solve, iteration count and whether the plain fp64
factorization took over; stall handling is forced from the host."""
import ctypes as C

import numpy as np
import pytest
import torch

import libmat_amd
from _util import backward_error, from_dev, stream, to_dev

pytestmark = pytest.mark.gpu


def ptr(t):
    return C.c_void_p(t.data_ptr())


def conditioned(n, kappa, seed, spd=False):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    if spd:
        lam = np.exp(rng.uniform(0, np.log(kappa), n))
        return (U * lam) @ U.T
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sig = np.geomspace(1.0, 1.0 / kappa, n)
    return (U * sig) @ V.T


def run(kind, A, B, passes, max_iter=30, pivoting=None):
    g = libmat_amd.load("f64")
    n, nr = B.shape
    opts = libmat_amd.IrOpts()
    g.lib.mat_gpu_ir_defaults(C.byref(opts), 1 if kind == "posv" else 0)
    opts.passes = passes
    opts.max_iter = max_iter
    if pivoting is not None:
        opts.pivoting = pivoting
    st = libmat_amd.IrStatus()
    dA, dB = to_dev(A, np.float64), to_dev(B, np.float64)
    dX = torch.zeros(n * nr, dtype=torch.float64, device="cuda")
    rc = getattr(g.lib, f"mat_gpu_{kind}_mixed")(stream(), n, nr, ptr(dA), n, ptr(dB), n, ptr(dX),
                                                 n, C.byref(opts), C.byref(st))
    g.check(rc, kind)
    torch.cuda.synchronize()
    return rc, from_dev(dX, n, nr), st


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("passes", [3, 6])
@pytest.mark.parametrize("pivoting", [0, 1, 2])
def test_gesv_mixed(kappa, passes, pivoting):
    n = 1538 if pivoting == 1 else 1536  # the butterfly path pads n to a multiple of 4
    A = conditioned(n, kappa, seed=int(np.log10(kappa)))
    B = np.random.default_rng(1).uniform(-1, 1, (n, 16))
    rc, X, st = run("gesv", A, B, passes, pivoting=pivoting)
    be = backward_error(A, X, B)
    Xref = np.linalg.solve(A, B)
    fe = np.abs(X - Xref).max() / np.abs(Xref).max()
    print(f"kappa={kappa:g} passes={passes} pivoting={pivoting} iters={st.iterations} "
          f"fell_back={st.fell_back} "
          f"berr={be:.2e} (lib {st.backward_error:.2e}) ferr={fe:.2e}")
    assert rc == 0
    assert be <= 1e-14 and st.backward_error <= 1e-14
    assert fe <= kappa * 1e-13
    if kappa <= 1e4 and passes == 6 and pivoting != 1:
        assert st.fell_back == 0


def conditioned_dev(n, kappa, seed, spd=False):
    """As conditioned(), generated on the GPU (column-major fp64 buffer as a
    flat tensor; the matrix is the transpose of the row-major product, which
    is a member of the same family)."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    U, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen))
    if spd:
        lam = torch.exp(torch.rand(n, dtype=torch.float64, device="cuda", generator=gen)
                        * np.log(kappa))
        A = (U * lam) @ U.T
        A = (A + A.T) / 2
    else:
        V, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen))
        sig = torch.logspace(0, -np.log10(kappa), n, dtype=torch.float64, device="cuda")
        A = (U * sig) @ V.T
    return A.contiguous().reshape(-1)


def run_dev(kind, dA, n, B, passes, pivoting=None):
    g = libmat_amd.load("f64")
    nr = B.shape[1]
    opts = libmat_amd.IrOpts()
    g.lib.mat_gpu_ir_defaults(C.byref(opts), 1 if kind == "posv" else 0)
    opts.passes = passes
    if pivoting is not None:
        opts.pivoting = pivoting
    st = libmat_amd.IrStatus()
    dB = to_dev(B, np.float64)
    dX = torch.zeros(n * nr, dtype=torch.float64, device="cuda")
    rc = getattr(g.lib, f"mat_gpu_{kind}_mixed")(stream(), n, nr, ptr(dA), n, ptr(dB), n, ptr(dX),
                                                 n, C.byref(opts), C.byref(st))
    g.check(rc, kind)
    torch.cuda.synchronize()
    A = dA.reshape(n, n).T  # logical column-major matrix
    X = dX.reshape(nr, n).T
    Bm = dB.reshape(nr, n).T
    R = A @ X - Bm
    be = float((R.abs().amax(0) / (A.abs().sum(1).max() * X.abs().amax(0) + Bm.abs().amax(0))).max())
    Xref = torch.linalg.solve(A, Bm)
    fe = float((X - Xref).abs().max() / Xref.abs().max())
    return rc, st, be, fe


@pytest.mark.parametrize("kappa", [1e2, 1e6])
@pytest.mark.parametrize("passes", [3, 6])
@pytest.mark.parametrize("pivoting", [0, 2])
def test_gesv_mixed_split_updates(kappa, passes, pivoting):
    """n = 4096: the large trailing updates run as split-BF16 MFMA GEMMs;
    partial and tournament pivoting."""
    n = 4096
    dA = conditioned_dev(n, kappa, seed=11)
    B = np.random.default_rng(5).uniform(-1, 1, (n, 16))
    rc, st, be, fe = run_dev("gesv", dA, n, B, passes, pivoting)
    print(f"n={n} kappa={kappa:g} passes={passes} pivoting={pivoting} iters={st.iterations} "
          f"fell_back={st.fell_back} berr={be:.2e} ferr={fe:.2e}")
    assert rc == 0 and be <= 1e-14 and fe <= kappa * 1e-13
    if passes == 6:
        assert st.fell_back == 0


@pytest.mark.parametrize("kappa", [1e2, 1e3])
def test_posv_mixed_fp8_updates(kappa):
    """n = 4096: the top-level trailing update runs in FP8 (tile-scaled); the
    ladder may refactor with 3-pass BF16 before the fp64 fallback."""
    n = 4096
    dA = conditioned_dev(n, kappa, seed=13, spd=True)
    B = np.random.default_rng(6).standard_normal((n, 4))
    rc, st, be, fe = run_dev("posv", dA, n, B, 8)
    print(f"fp8 n={n} kappa={kappa:g} iters={st.iterations} precision={st.precision} "
          f"fell_back={st.fell_back} berr={be:.2e} ferr={fe:.2e}")
    assert rc == 0 and be <= 1e-14 and fe <= kappa * 1e-13
    assert st.precision in (8, 3)
    if kappa <= 1e2:
        assert st.fell_back == 0


@pytest.mark.parametrize("kind", ["gesv", "posv"])
def test_mixed_nan_is_not_converged(kind):
    """A NaN in the data must never pass the backward-error test: the norms
    propagate NaN, refinement stops and the fp64 solver's result is reported."""
    n = 300
    A = conditioned(n, 1e2, seed=4, spd=(kind == "posv"))
    A[5, 7] = np.nan
    B = np.random.default_rng(8).uniform(-1, 1, (n, 2))
    g = libmat_amd.load("f64")
    opts = libmat_amd.IrOpts()
    g.lib.mat_gpu_ir_defaults(C.byref(opts), 1 if kind == "posv" else 0)
    st = libmat_amd.IrStatus()
    dA, dB = to_dev(A, np.float64), to_dev(B, np.float64)
    dX = torch.zeros(n * 2, dtype=torch.float64, device="cuda")
    getattr(g.lib, f"mat_gpu_{kind}_mixed")(stream(), n, 2, ptr(dA), n, ptr(dB), n, ptr(dX), n,
                                            C.byref(opts), C.byref(st))
    torch.cuda.synchronize()
    assert not (st.backward_error <= 1e-14)
    if kind == "gesv":
        assert st.fell_back == 1


def test_gesv_mixed_forced_stall_falls_back():
    n = 512
    A = conditioned(n, 1e6, seed=3)
    B = np.random.default_rng(2).uniform(-1, 1, (n, 4))
    rc, X, st = run("gesv", A, B, passes=1, max_iter=0)
    assert rc == 0 and st.fell_back == 1 and st.iterations == 0
    assert backward_error(A, X, B) <= 1e-14


@pytest.mark.parametrize("kappa", [1e2, 1e3])
def test_posv_mixed_fp8(kappa):
    n = 1536
    A = conditioned(n, kappa, seed=7, spd=True)
    B = np.random.default_rng(3).standard_normal((n, 4))
    rc, X, st = run("posv", A, B, passes=8)
    be = backward_error(A, X, B)
    Xref = np.linalg.solve(A, B)
    fe = np.abs(X - Xref).max() / np.abs(Xref).max()
    print(f"fp8 kappa={kappa:g} iters={st.iterations} fell_back={st.fell_back} berr={be:.2e} "
          f"ferr={fe:.2e}")
    assert rc == 0 and be <= 1e-14 and fe <= kappa * 1e-13


def test_posv_mixed_not_spd():
    n = 256
    A = conditioned(n, 10, seed=9, spd=True)
    A[100, 100] = -1.0
    B = np.ones((n, 1))
    rc, X, st = run("posv", A, B, passes=8)
    assert rc == -1
