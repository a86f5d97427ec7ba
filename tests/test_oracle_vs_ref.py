"""Oracle vs the reference compiled from its own sources (oracle/_ref):
bit-exact on seeded inputs across sizes that cross the 24-column PLU
blocks and the Cholesky sizes.  Skipped where oracle/_ref was not built."""
import numpy as np
import pytest

from oracle.oracle import Oracle, available

pytestmark = pytest.mark.skipif(not available("f32", "ref"), reason="oracle/_ref not built")

SIZES = [1, 2, 3, 5, 23, 24, 25, 47, 48, 49, 64, 65, 97, 130]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_bit_exact(prec):
    o, r = Oracle(prec), Oracle(prec, "ref")
    rng = np.random.default_rng(7)
    dt = o.dtype
    for n in SIZES:
        A = rng.uniform(-1, 1, (n, n)).astype(dt)
        B = rng.uniform(-1, 1, (n, n + 3)).astype(dt)
        Cm = rng.uniform(-1, 1, (n, n + 3)).astype(dt)
        b = rng.uniform(-1, 1, n).astype(dt)
        assert np.array_equal(o.gemm(0.5, A, B, 2.0, Cm), r.gemm(0.5, A, B, 2.0, Cm))
        for x, y in zip(o.plu(A), r.plu(A)):
            assert np.array_equal(x, y), n
        for x, y in zip(o.lu(A), r.lu(A)):
            assert np.array_equal(x, y), n
        S = (A @ A.T + n * np.eye(n)).astype(dt)
        for x, y in zip(o.chol(S), r.chol(S)):
            assert np.array_equal(x, y)
        assert np.array_equal(o.solve(A, b), r.solve(A, b))
        for x, y in zip(o.solve_spd(S, b), r.solve_spd(S, b)):
            assert np.array_equal(x, y)
        T = (np.tril(A) + 3 * np.eye(n)).astype(dt)
        for f in ("solve_tril", "solve_tril_unit", "solve_trilt"):
            assert np.array_equal(getattr(o, f)(T, b), getattr(r, f)(T, b)), f
        assert np.array_equal(o.solve_triu(T.T, b), r.solve_triu(T.T, b))
        Bm = rng.uniform(-1, 1, (n, 4)).astype(dt)
        assert np.array_equal(o.solve_multi(A, Bm), r.solve_multi(A, Bm))
        assert o.det(A) == r.det(A) or np.isnan(o.det(A))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", [False, True])
def test_syrk_bit_exact(prec, uplo, trans):
    o, r = Oracle(prec), Oracle(prec, "ref")
    rng = np.random.default_rng(11)
    dt = o.dtype
    for n, k in ((1, 1), (2, 3), (5, 1), (24, 7), (49, 33), (65, 130)):
        A = rng.uniform(-1, 1, (k, n) if trans else (n, k)).astype(dt)
        Cm = rng.uniform(-1, 1, (n, n)).astype(dt)
        for alpha, beta in ((1.0, 0.0), (-0.75, 1.5)):
            assert np.array_equal(o.syrk(alpha, A, beta, Cm, uplo, trans),
                                  r.syrk(alpha, A, beta, Cm, uplo, trans)), (n, k)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_qr_bit_exact(prec):
    """mat_qr / mat_qr_r (mat.h:3907-4192): the unblocked path (n < 64) and
    the blocked compact-WY path (n >= 64, blocks of 16, a ragged last
    block), tall and square, and a zero column (no reflector)."""
    o, r = Oracle(prec), Oracle(prec, "ref")
    rng = np.random.default_rng(12)
    for m, n in ((1, 1), (3, 3), (4, 3), (10, 7), (63, 63), (64, 64), (100, 70), (97, 96),
                 (150, 130)):
        A = rng.uniform(-1, 1, (m, n)).astype(o.dtype)
        if n > 5:
            A[:, 5] = 0
        Qo, Ro = o.qr(A)
        Qr, Rr = r.qr(A)
        assert np.array_equal(Qo, Qr) and np.array_equal(Ro, Rr), (m, n)
        assert np.array_equal(o.qr_r(A), r.qr_r(A)), (m, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_inv_and_multi_rhs_bit_exact(prec):
    """mat_inv (mat.h:6166-6225), the multi-RHS Cholesky solve
    (mat.h:4854-4865 per column) and the per-matrix batch loop."""
    o, r = Oracle(prec), Oracle(prec, "ref")
    rng = np.random.default_rng(13)
    dt = o.dtype
    for n in SIZES:
        A = rng.uniform(-1, 1, (n, n)).astype(dt)
        assert np.array_equal(o.inv(A), r.inv(A)), n
        S = (A @ A.T + n * np.eye(n)).astype(dt)
        B = rng.uniform(-1, 1, (n, 3)).astype(dt)
        Xo, ro = o.chol_solve_multi(S, B)
        Xr, rr = r.chol_solve_multi(S, B)
        assert ro == rr == 0 and np.array_equal(Xo, Xr), n
    # a batch with general, SPD and one non-SPD member
    ns = np.array([1, 3, 7, 24, 25, 5], np.int32)
    kinds = np.array([0, 1, 0, 1, 0, 1], np.int8)
    mats, vecs = [], []
    for n, k in zip(ns, kinds):
        M = rng.uniform(-1, 1, (n, n))
        M = M + n * np.eye(n) if k == 0 else M @ M.T + np.eye(n)
        mats.append(M.astype(dt))
        vecs.append(rng.uniform(-1, 1, n).astype(dt))
    mats[-1] = -mats[-1]
    a_off = np.concatenate([[0], np.cumsum(ns.astype(np.int64) ** 2)[:-1]])
    v_off = np.concatenate([[0], np.cumsum(ns.astype(np.int64))[:-1]])
    Ap = np.concatenate([m.ravel(order="F") for m in mats])
    bp = np.concatenate(vecs)
    xo, io = o.batch_solve(ns, kinds, Ap, a_off, bp, v_off)
    xr, ir = r.batch_solve(ns, kinds, Ap, a_off, bp, v_off)
    assert np.array_equal(xo, xr) and np.array_equal(io, ir)
    assert io.tolist() == [0, 0, 0, 0, 0, -1]
