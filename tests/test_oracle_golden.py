"""The oracle pinned: bit-exact against the fixtures the reference wrote,
against the reference's own known-answer tests (tests/test_mat_*.c), and
against invariants / library routines that a plausible mistake would break."""
import numpy as np
import pytest
import scipy.linalg as sla

from _util import DT, EPS, apply_row_perm, golden, rel_fro
from oracle.oracle import Oracle

PRECS = ["f32", "f64"]


@pytest.fixture(scope="module", params=PRECS)
def orc(request):
    return Oracle(request.param)


def keys(g, prefix):
    return sorted({k.rsplit("_", 1)[0] for k in g if k.startswith(prefix)})


# ---------------------------------------------------- fixtures, bit-exact --
def test_gemm_fixtures(orc):
    g = golden(orc.prec)
    for t in [k[:-2] for k in g if k.startswith("gemm_") and k.endswith("_A")]:
        out = orc.gemm(0.7, g[t + "_A"], g[t + "_B"], 0.3, g[t + "_C0"])
        assert np.array_equal(out, g[t + "_C"]), t


def test_cfg1_fixture():
    o, g = Oracle("f32"), golden("f32")
    assert np.array_equal(o.mul(g["cfg1_A"], g["cfg1_B"]), g["cfg1_C"])
    L, U, p, s = o.plu(g["cfg1_A"])
    assert np.array_equal(L, g["cfg1_L"]) and np.array_equal(U, g["cfg1_U"])
    assert np.array_equal(p, g["cfg1_p"]) and s == g["cfg1_swaps"]
    assert np.array_equal(o.solve(g["cfg1_A"], g["cfg1_b"]), g["cfg1_x"])


def test_plu_solve_det_fixtures(orc):
    g = golden(orc.prec)
    tags = [k[:-2] for k in g if (k.startswith("plu_") or k.startswith("sing") or
                                  k.startswith("zerocol") or k.startswith("rank2")) and k.endswith("_A")]
    assert len(tags) >= 15
    for t in tags:
        L, U, p, s = orc.plu(g[t + "_A"])
        assert np.array_equal(L, g[t + "_L"]) and np.array_equal(U, g[t + "_U"]), t
        assert np.array_equal(p, g[t + "_p"]) and s == g[t + "_swaps"], t
        assert orc.det(g[t + "_A"]) == g[t + "_det"] or (np.isnan(g[t + "_det"])), t
        if t + "_x" in g:
            assert np.array_equal(orc.solve(g[t + "_A"], g[t + "_b"]), g[t + "_x"]), t


def test_lu_fixtures(orc):
    g = golden(orc.prec)
    for t in [k[:-2] for k in g if k.startswith("lu_") and k.endswith("_A")]:
        L, U, p, q, s = orc.lu(g[t + "_A"])
        for a, b in ((L, "_L"), (U, "_U"), (p, "_p"), (q, "_q")):
            assert np.array_equal(a, g[t + b]), t + b
        assert s == g[t + "_swaps"]


def test_chol_fixtures(orc):
    g = golden(orc.prec)
    for t in [k[:-2] for k in g if k.startswith("chol_") and k.endswith("_A")]:
        L, rc = orc.chol(g[t + "_A"])
        assert rc == g[t + "_rc"] == 0 and np.array_equal(L, g[t + "_L"]), t
        x, rs = orc.solve_spd(g[t + "_A"], g[t + "_b"])
        assert rs == 0 and np.array_equal(x, g[t + "_x"]), t
    for t in ("notpd2", "indef40"):
        assert orc.chol(g[t + "_A"])[1] == -1 == g[t + "_rc"]
        assert orc.solve_spd(g[t + "_A"], np.ones(g[t + "_A"].shape[0]))[1] == -1


def test_tri_fixtures(orc):
    g = golden(orc.prec)
    for t in [k[:-2] for k in g if k.startswith("tri_") and k.endswith("_T")]:
        T, b = g[t + "_T"], g[t + "_b"]
        assert np.array_equal(orc.solve_tril(np.tril(T), b), g[t + "_tril"])
        assert np.array_equal(orc.solve_tril_unit(np.tril(T), b), g[t + "_tril_unit"])
        assert np.array_equal(orc.solve_triu(np.triu(T), b), g[t + "_triu"])
        assert np.array_equal(orc.solve_trilt(np.tril(T), b), g[t + "_trilt"])


# ------------------------------------------ reference known-answer tests --
def test_mul_known_answers(orc):
    # tests/test_mat_mul.c:6-90
    A = np.array([[1, 2], [3, 4]])
    assert np.array_equal(orc.mul(A, [[5, 6], [7, 8]]), [[19, 22], [43, 50]])
    assert np.array_equal(orc.mul(A, np.eye(2)), A)
    assert np.array_equal(orc.mul([[1, 2, 3], [4, 5, 6]], [[7, 8], [9, 10], [11, 12]]),
                          [[58, 64], [139, 154]])
    assert np.array_equal(orc.mul(A, [[5], [6]]), [[17], [39]])


def test_plu_known_answers(orc):
    # tests/test_mat_plu.c:17-80, 121-155
    for A in ([[4, 3], [6, 3]], [[1, 2, 3], [0, 1, 4], [5, 6, 0]]):
        L, U, p, s = orc.plu(A)
        assert np.abs(apply_row_perm(A, p) - L @ U).max() <= 1e-5
    A = np.array([[6, 1, 1], [4, -2, 5], [2, 8, 7]])
    L, U, p, q, s = orc.lu(A)
    det_lu = (1 if s % 2 == 0 else -1) * np.prod(np.diag(U))
    assert abs(orc.det(A) - det_lu) < 1e-4
    assert abs(orc.det(A) - np.linalg.det(A)) < 1e-3


def test_solve_known_answers(orc):
    # tests/test_mat_solve.c:7-136
    assert np.allclose(orc.solve([[5]], [15]), [3], atol=1e-6)
    assert np.allclose(orc.solve([[2, 3], [4, 5]], [8, 14]), [1, 2], atol=1e-5)
    assert np.allclose(orc.solve([[1, 2, 3], [0, 1, 4], [5, 6, 0]], [14, 14, 17]), [1, 2, 3],
                       atol=1e-4)


def test_spd_known_answers(orc):
    # tests/test_mat_chol.c:67-131, test_mat_solve_spd.c:50-122
    L, rc = orc.chol([[4, 2], [2, 5]])
    assert rc == 0 and np.allclose(L, [[2, 0], [1, 2]], atol=1e-6)
    A = np.array([[4, 12, -16], [12, 37, -43], [-16, -43, 98]], float)
    L, rc = orc.chol(A)
    assert rc == 0 and np.allclose(L, [[2, 0, 0], [6, 1, 0], [-8, 5, 3]], atol=1e-5)
    x, rc = orc.solve_spd([[4, 2], [2, 3]], [8, 8])
    assert rc == 0 and np.allclose(x, [1, 2], atol=1e-5)
    x, rc = orc.solve_spd([[5, 4, 1], [4, 6, 4], [1, 4, 5]], [16, 28, 24])
    assert rc == 0 and np.allclose(x, [1, 2, 3], atol=1e-4)
    assert orc.solve_spd([[1, 2], [2, 1]], [1, 1])[1] == -1
    assert orc.chol([[1, 2], [2, 1]])[1] == -1


def test_tri_known_answers(orc):
    # tests/test_mat_solve_tri.c:7-123
    assert np.allclose(orc.solve_tril([[2, 0], [3, 4]], [2, 11]), [1, 2])
    assert np.allclose(orc.solve_tril_unit([[1, 0], [3, 1]], [1, 5]), [1, 2])
    assert np.allclose(orc.solve_triu([[2, 3], [0, 4]], [8, 8]), [1, 2])
    assert np.allclose(orc.solve_trilt([[2, 0], [3, 4]], [8, 8]), [1, 2])


def test_det_singular(orc):
    # tests/test_mat_det.c:55-71: the near-zero pivot is skipped, det -> 0
    assert abs(orc.det([[1, 2, 3], [4, 5, 6], [5, 7, 9]])) < 1e-5


# ----------------------------------------- mathematics pins the rest ------
@pytest.mark.parametrize("n", [5, 37, 100])
def test_plu_invariants_and_lapack(orc, n):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n)).astype(DT[orc.prec])
    L, U, p, s = orc.plu(A)
    assert sorted(p.tolist()) == list(range(n))
    assert np.allclose(np.triu(L, 1), 0) and np.allclose(np.diag(L), 1)
    assert np.allclose(np.tril(U, -1), 0) and np.abs(L).max() <= 1
    assert rel_fro(L @ U, apply_row_perm(A, p)) <= 4 * n * EPS[orc.prec]
    # LAPACK getrf picks the same pivot rule (first max |a|); compare the
    # permutation (well-separated random data) and the sign of det P
    lu, piv = sla.lu_factor(A.astype(np.float64))
    perm = np.arange(n)
    for k, r in enumerate(piv):
        perm[[k, r]] = perm[[r, k]]
    assert np.array_equal(perm, p)
    assert (-1) ** s == np.linalg.det(np.eye(n)[p])


def test_gemm_brute_force(orc):
    rng = np.random.default_rng(3)
    A, B, Cm = (rng.uniform(-1, 1, s) for s in ((31, 45), (45, 12), (31, 12)))
    out = orc.gemm(1.5, A, B, -0.5, Cm)
    ref = 1.5 * A @ B - 0.5 * Cm
    assert rel_fro(out, ref) <= 45 * EPS[orc.prec]
    # beta == 0 must not read C (the reference zero-fills, mat.h:994-998)
    nanC = np.full((31, 12), np.nan)
    assert np.all(np.isfinite(orc.gemm(1.0, A, B, 0.0, nanC)))


@pytest.mark.parametrize("n", [6, 50])
def test_chol_vs_lapack(orc, n):
    rng = np.random.default_rng(n + 1)
    g = rng.standard_normal((n, n))
    A = (g @ g.T / n + np.eye(n)).astype(DT[orc.prec])
    L, rc = orc.chol(A)
    assert rc == 0
    assert rel_fro(L, np.linalg.cholesky(A.astype(np.float64))) <= 8 * n * EPS[orc.prec]
    # only the lower triangle is read (mat.h:6072-6082)
    junk = A + np.triu(np.full((n, n), 7.0), 1)
    assert np.array_equal(orc.chol(junk)[0], L)


def test_lu_full_pivot_invariants(orc):
    rng = np.random.default_rng(11)
    A = rng.uniform(-10, 10, (30, 30)).astype(DT[orc.prec])
    L, U, p, q, s = orc.lu(A)
    PAQ = A[p][:, q]
    assert rel_fro(L @ U, PAQ) <= 200 * EPS[orc.prec]
    # full pivoting: |U[k,k]| is the max of the trailing block at step k
    assert np.all(np.abs(np.diag(U))[:-1] >= np.abs(np.triu(U, 1)).max(axis=1)[:-1] * (1 - 1e-6))


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", [False, True])
def test_syrk_brute_force_and_known_answer(orc, uplo, trans):
    """mat_syrk / mat_syrk_t (mat.h:5641-5734): the updated triangle against
    an fp64 brute force, the other triangle untouched bit for bit, and a 2x2
    known answer: A = [1 2; 3 4] gives A A^T = [5 11; 11 25] and
    A^T A = [10 14; 14 20]."""
    rng = np.random.default_rng(5)
    n, k = 37, 19
    A = rng.uniform(-1, 1, (k, n) if trans else (n, k)).astype(orc.dtype)
    Cm = rng.uniform(-1, 1, (n, n)).astype(orc.dtype)
    out = orc.syrk(-0.5, A, 2.0, Cm, uplo, trans)
    A64 = A.astype(np.float64)
    full = -0.5 * (A64.T @ A64 if trans else A64 @ A64.T) + 2.0 * Cm.astype(np.float64)
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    eps = np.finfo(orc.dtype).eps
    assert np.abs(out[tri] - full[tri]).max() <= 4 * k * eps * (np.abs(A64).max() ** 2 * k + 2)
    assert np.array_equal(out[~tri], Cm[~tri])
    a = np.array([[1.0, 2.0], [3.0, 4.0]], dtype=orc.dtype)
    want = np.array([[10.0, 14.0], [14.0, 20.0]]) if trans else np.array([[5.0, 11.0],
                                                                            [11.0, 25.0]])
    got = orc.syrk(1.0, a, 0.0, np.full((2, 2), 7.0, dtype=orc.dtype), uplo, trans)
    keep = np.full((2, 2), 7.0)
    expect = np.where(np.tril(np.ones((2, 2), bool)) if uplo == "L"
                      else np.triu(np.ones((2, 2), bool)), want, keep)
    assert np.array_equal(got, expect)


def _qr_invariants(Q, R, A, eps):
    m, n = A.shape
    assert np.abs(Q.T @ Q - np.eye(m)).max() <= 20 * m * eps
    assert np.abs(Q @ np.triu(R) - A).max() <= 20 * m * eps * max(1.0, np.abs(A).max())
    # the reference's own is_upper_triangular tolerance (rows * eps_default)
    assert np.abs(np.tril(R, -1)).max() <= m * (1e-6 if eps > 1e-10 else 1e-9) * max(
        1.0, np.abs(A).max())


def test_qr_known_answers_and_lapack(orc):
    """The oracle's QR against what fixes it independently:
    - the reference's known-answer inputs (tests/test_mat_qr.c:32-120): the
      textbook 3x3 (|R| = [[14,21,14],[0,175,70],[0,0,35]]), the 4x3
      overdetermined case, the identity;
    - LAPACK's Householder QR (numpy), which uses the same sign rule
      (beta = -sign(x0) ||x||), so R and the full Q agree to rounding for
      both paths (unblocked n < 64 and blocked n >= 64);
    - the identity, where libmat reflects every column (Q = R = -I) and
      LAPACK does not (its tau is 0 when the tail is zero; the same rule
      flips the last row of R of a square matrix)."""
    dt = orc.dtype
    eps = EPS[orc.prec]
    A = np.array([[12, -51, 4], [6, 167, -68], [-4, 24, -41]], dt)
    Q, R = orc.qr(A)
    _qr_invariants(Q, R, A, eps)
    want = np.array([[14, 21, 14], [0, 175, 70], [0, 0, 35]], np.float64)
    assert np.abs(np.abs(np.triu(R)) - want).max() <= 1e3 * eps
    A = np.arange(1, 13, dtype=dt).reshape(4, 3)
    Q, R = orc.qr(A)
    _qr_invariants(Q, R, A, eps)
    Q, R = orc.qr(np.eye(3, dtype=dt))
    assert np.array_equal(Q, -np.eye(3, dtype=dt)) and np.array_equal(np.triu(R), -np.eye(3, dtype=dt))
    rng = np.random.default_rng(21)
    for m, n in ((5, 5), (40, 25), (64, 64), (150, 100), (130, 130)):
        A = rng.uniform(-1, 1, (m, n)).astype(dt)
        Q, R = orc.qr(A)
        _qr_invariants(Q, R, A, eps)
        Ql, Rl = np.linalg.qr(A.astype(np.float64), mode="complete")
        if m == n:
            # a square matrix's last column is a one-element reflector:
            # libmat reflects it (a sign flip), LAPACK takes tau = 0
            Rl[-1, :] *= -1
            Ql[:, -1] *= -1
        kappa = np.linalg.cond(A.astype(np.float64))
        tol = 8 * kappa * np.sqrt(m) * eps
        assert rel_fro(np.triu(R), Rl) <= tol, (m, n)
        assert rel_fro(Q, Ql) <= tol, (m, n)
        assert np.array_equal(orc.qr_r(A), R), (m, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_inv_closed_form(prec):
    """mat_inv (mat.h:6166-6225) seeds its right-hand sides with the
    identity permuted by p^-1 (column i holds a one at row p[i]) where
    A^-1 = U^-1 L^-1 P needs P, so it returns A^-1 (P^T)^2: A^-1 itself
    exactly when the row permutation is its own inverse (no swaps, or
    disjoint transpositions), as for the reference's own diagonally
    dominant test matrices (tests/test_mat_inv.c:82-118).  Pinned here
    against numpy's inverse (LAPACK getri) in both cases."""
    o = Oracle(prec)
    rng = np.random.default_rng(21)
    tol = 1e-3 if prec == "f32" else 1e-10
    for n in (1, 2, 3, 8, 25, 60):
        A = rng.uniform(-1, 1, (n, n)).astype(o.dtype)
        X = o.inv(A).astype(np.float64)
        _, _, p, _ = o.plu(A)
        Pt = np.zeros((n, n))
        Pt[np.asarray(p), np.arange(n)] = 1.0  # column i: a one at row p[i]
        Ainv = np.linalg.inv(A.astype(np.float64))
        want = Ainv @ Pt @ Pt
        assert np.abs(X - want).max() <= tol * max(1.0, np.abs(want).max()), n
        D = (A + n * np.eye(n)).astype(o.dtype)  # diagonally dominant: no swaps
        XD = o.inv(D).astype(np.float64)
        assert np.abs(D.astype(np.float64) @ XD - np.eye(n)).max() <= tol, n
