"""Drop-in libmat entry points on the GPU vs the reference: committed
fixtures written by the reference, its own known-answer and random tests
(tests/test_mat_{mul,plu,lu,chol,solve,solve_spd,solve_tri,det,inv}.c),
and the north-star tolerances: GEMM relative Frobenius error <= c*K*eps,
pivots identical except at near-ties, backward errors at the working
precision."""
import numpy as np
import pytest

from _util import DT, EPS, apply_row_perm, backward_error, golden, pivots_match, rel_fro

pytestmark = pytest.mark.gpu
PRECS = ["f32", "f64"]


@pytest.mark.parametrize("prec", PRECS)
def test_mul_known_answers(gpu, prec):
    g = gpu[prec]
    A = np.array([[1, 2], [3, 4]])
    assert np.abs(g.mat_mul(A, [[5, 6], [7, 8]]) - [[19, 22], [43, 50]]).max() <= 1e-6
    assert np.abs(g.mat_mul(A, np.eye(2)) - A).max() <= 1e-6
    assert np.abs(g.mat_mul([[1, 2, 3], [4, 5, 6]], [[7, 8], [9, 10], [11, 12]])
                  - [[58, 64], [139, 154]]).max() <= 1e-6
    assert np.abs(g.mat_rmul(A, [[5], [6]]) - [[17], [39]]).max() <= 1e-6


@pytest.mark.parametrize("prec", PRECS)
def test_gemm_fixtures(gpu, prec):
    g, gold = gpu[prec], golden(prec)
    for t in [k[:-2] for k in gold if k.startswith("gemm_") and k.endswith("_A")]:
        K = gold[t + "_A"].shape[1]
        out = g.mat_gemm(0.7, gold[t + "_A"], gold[t + "_B"], 0.3, gold[t + "_C0"])
        assert rel_fro(out, gold[t + "_C"]) <= 2 * K * EPS[prec], t


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", [False, True])
def test_syrk_against_oracle(gpu, prec, uplo, trans):
    """mat_syrk / mat_syrk_t (mat.h:5670-5677, 5727-5734) against the oracle
    on the updated triangle (|error| <= 2 K eps of the row norms), the other
    triangle bit for bit untouched; shapes cover both the 64x64 and the
    128x128 GEMM tiles and tiles that straddle the diagonal."""
    from oracle.oracle import Oracle
    g, o = gpu[prec], Oracle(prec)
    rng = np.random.default_rng(3)
    for n, k in ((1, 1), (3, 5), (64, 64), (130, 33), (257, 300), (2100, 64)):
        A = rng.uniform(-1, 1, (k, n) if trans else (n, k)).astype(DT[prec])
        Cm = rng.uniform(-1, 1, (n, n)).astype(DT[prec])
        for alpha, beta in ((1.0, 0.0), (-0.75, 1.5)):
            got = g.mat_syrk(alpha, A, beta, Cm, uplo, trans)
            want = o.syrk(alpha, A, beta, Cm, uplo, trans)
            tri = (np.tril(np.ones((n, n), bool)) if uplo == "L"
                   else np.triu(np.ones((n, n), bool)))
            scale = abs(alpha) * k + abs(beta)
            assert np.abs(got[tri] - want[tri]).max() <= 2 * k * EPS[prec] * scale, (n, k)
            assert np.array_equal(got[~tri], Cm[~tri]), (n, k)


def test_cfg1_against_reference(gpu):
    g, gold = gpu["f32"], golden("f32")
    C = g.mat_mul(gold["cfg1_A"], gold["cfg1_B"])
    assert rel_fro(C, gold["cfg1_C"]) <= 256 * EPS["f32"]
    L, U, p, s = g.mat_plu(gold["cfg1_A"])
    ok, k = pivots_match(p, gold["cfg1_p"], gold["cfg1_L"], "f32")
    assert ok, f"pivot mismatch at step {k}"
    if k is None:
        assert s == gold["cfg1_swaps"]
        assert rel_fro(L, gold["cfg1_L"]) <= 1e-4 and rel_fro(U, gold["cfg1_U"]) <= 1e-4
    assert rel_fro(L @ U, apply_row_perm(gold["cfg1_A"], p)) <= 256 * 4 * EPS["f32"]
    x = g.mat_solve(gold["cfg1_A"], gold["cfg1_b"])
    assert backward_error(gold["cfg1_A"], x, gold["cfg1_b"]) <= 256 * EPS["f32"]
    assert rel_fro(x, gold["cfg1_x"]) <= 1e-3


@pytest.mark.parametrize("prec", PRECS)
def test_plu_fixtures(gpu, prec):
    g, gold = gpu[prec], golden(prec)
    for t in [k[:-2] for k in gold if k.startswith("plu_") and k.endswith("_A")]:
        A = gold[t + "_A"]
        n = A.shape[0]
        L, U, p, s = g.mat_plu(A)
        ok, k = pivots_match(p, gold[t + "_p"], gold[t + "_L"], prec)
        assert ok, f"{t}: pivot mismatch at step {k}"
        if k is None:
            assert s == gold[t + "_swaps"], t
            assert rel_fro(U, gold[t + "_U"]) <= 64 * n * EPS[prec], t
        assert np.allclose(np.triu(L, 1), 0) and np.allclose(np.diag(L), 1)
        assert np.allclose(np.tril(U, -1), 0)
        assert rel_fro(L @ U, apply_row_perm(A, p)) <= 4 * n * EPS[prec], t
        x = g.mat_solve(A, gold[t + "_b"])
        assert backward_error(A, x, gold[t + "_b"]) <= 8 * n * EPS[prec], t
        det, dref = g.mat_det(A), float(gold[t + "_det"])
        if np.isfinite(dref):  # det of a 100x100 uniform matrix overflows float
            assert abs(det - dref) <= 1e-3 * max(1.0, abs(dref)), t
        else:
            assert det == dref, t


@pytest.mark.parametrize("prec", PRECS)
def test_singular_eps_skip(gpu, prec):
    """Near-zero pivots are swapped in but not divided by (mat.h:4234)."""
    g, gold = gpu[prec], golden(prec)
    # exact zero column: same pivots as the reference, det exactly 0
    A = gold["zerocol30_A"]
    L, U, p, s = g.mat_plu(A)
    ok, k = pivots_match(p, gold["zerocol30_p"], gold["zerocol30_L"], prec)
    assert ok, k
    assert U[7, 7] == 0 and g.mat_det(A) == 0
    assert np.all(np.isfinite(L)) and np.all(np.isfinite(U))
    assert rel_fro(L @ U, apply_row_perm(A, p)) <= 200 * EPS[prec]
    # tests/test_mat_det.c:55-71
    assert abs(g.mat_det(gold["sing3_A"])) < 1e-5
    # rank 2: the first two pivots are the reference's, the rest is rounding noise
    A = gold["rank2_40_A"]
    L, U, p, s = g.mat_plu(A)
    assert np.array_equal(p[:2], gold["rank2_40_p"][:2])
    assert np.all(np.isfinite(U)) and abs(g.mat_det(A)) < 1e-6


@pytest.mark.parametrize("prec", PRECS)
def test_lu_full_pivot_fixtures(gpu, prec):
    g, gold = gpu[prec], golden(prec)
    for t in [k[:-2] for k in gold if k.startswith("lu_") and k.endswith("_A")]:
        A = gold[t + "_A"]
        n = A.shape[0]
        L, U, p, q, s = g.mat_lu(A)
        assert np.array_equal(p, gold[t + "_p"]) and np.array_equal(q, gold[t + "_q"]), t
        assert s == gold[t + "_swaps"], t
        assert rel_fro(L, gold[t + "_L"]) <= 32 * n * EPS[prec], t
        assert rel_fro(U, gold[t + "_U"]) <= 32 * n * EPS[prec], t


@pytest.mark.parametrize("prec", PRECS)
def test_chol_and_spd_fixtures(gpu, prec):
    g, gold = gpu[prec], golden(prec)
    for t in [k[:-2] for k in gold if k.startswith("chol_") and k.endswith("_A")]:
        A = gold[t + "_A"]
        n = A.shape[0]
        L, rc = g.mat_chol(A)
        assert rc == 0, t
        assert np.all(np.triu(L, 1) == 0)
        assert rel_fro(L, gold[t + "_L"]) <= 16 * n * EPS[prec], t
        x, rs = g.mat_solve_spd(A, gold[t + "_b"])
        assert rs == 0 and rel_fro(x, gold[t + "_x"]) <= 1e3 * n * EPS[prec], t
    for t in ("notpd2", "indef40"):
        A = gold[t + "_A"]
        assert g.mat_chol(A)[1] == -1
        x, rs = g.mat_solve_spd(A, np.ones(A.shape[0]))
        assert rs == -1 and np.all(x == 0)  # x untouched (mat.h:4855-4859)


@pytest.mark.parametrize("prec", PRECS)
def test_triangular_fixtures(gpu, prec):
    g, gold = gpu[prec], golden(prec)
    for t in [k[:-2] for k in gold if k.startswith("tri_") and k.endswith("_T")]:
        T, b = gold[t + "_T"], gold[t + "_b"]
        n = T.shape[0]
        tol = 64 * n * EPS[prec]
        Lt, Ut = np.tril(T), np.triu(T)
        Lu = np.tril(T, -1) + np.eye(n)
        cases = [("tril", g.mat_solve_tril, Lt, Lt), ("tril_unit", g.mat_solve_tril_unit, Lt, Lu),
                 ("triu", g.mat_solve_triu, Ut, Ut), ("trilt", g.mat_solve_trilt, Lt, Lt.T)]
        for name, fn, arg, op in cases:
            x = fn(arg, b)
            assert backward_error(op, x, b) <= 4 * n * EPS[prec], (t, name)
            if name != "tril_unit":  # random unit-lower systems are exponentially ill-conditioned
                assert rel_fro(x, gold[t + "_" + name]) <= tol, (t, name)


def _ref_rand(rng, n):
    # (rand() % 100) / 10 - 5 as in tests/test_mat_plu.c:87, from a seeded generator
    return rng.integers(0, 100, (n, n)) / 10.0 - 5.0


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n,tol", [(10, 1e-4), (50, 1e-4), (100, 1e-3)])
def test_reference_random_plu_solve(gpu, prec, n, tol):
    """tests/test_mat_plu.c:82-119 and tests/test_mat_solve.c:79-118."""
    g = gpu[prec]
    rng = np.random.default_rng(42 + n)
    A = _ref_rand(rng, n)
    A[np.diag_indices(n)] = np.abs(A).sum(axis=1) - np.abs(np.diag(A)) + 1.0
    A = A.astype(DT[prec])
    L, U, p, s = g.mat_plu(A)
    assert np.abs(apply_row_perm(A, p) - L @ U).max() < tol
    b = (rng.integers(0, 100, n) / 10.0).astype(DT[prec])
    x = g.mat_solve(A, b)
    assert np.abs(A.astype(np.float64) @ x - b).max() < {10: 1e-4, 50: 1e-3, 100: 1e-2}[n]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n,tol", [(10, 1e-4), (50, 1e-3), (100, 1e-2), (128, 1e-2), (200, 5e-2)])
def test_reference_random_chol(gpu, prec, n, tol):
    """tests/test_mat_chol.c:7-30 (B B^T + n I) and the L L^T = A check."""
    g = gpu[prec]
    rng = np.random.default_rng(7 + n)
    B = _ref_rand(rng, n)
    A = (B @ B.T + n * np.eye(n)).astype(DT[prec])
    L, rc = g.mat_chol(A)
    assert rc == 0 and np.all(np.diag(L) > 0)
    assert np.abs(L.astype(np.float64) @ L.T - A).max() < tol * max(1, np.abs(A).max() / 100)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n", [70, 130, 200, 256])
def test_det_inv(gpu, prec, n):
    """mat_inv solves with n right-hand sides: n = 70 runs the 65..128-row
    split triangular leaves, 130..256 the two-leaf path, lower unit and
    upper non-unit."""
    g = gpu[prec]
    A = np.array([[2, 0, 0, 0], [0, 3, 0, 0], [0, 0, 4, 0], [0, 0, 0, 3]], DT[prec])
    assert abs(g.mat_det(A) - 72) < 1e-4  # tests/test_mat_det.c:73-89
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
    Ai = g.mat_inv(A.astype(DT[prec]))
    assert np.abs(A @ Ai - np.eye(n)).max() < 100 * n * EPS[prec]


@pytest.mark.parametrize("prec", PRECS)
def test_one_by_one(gpu, prec):
    """1 x 1 systems against the oracle: a regular pivot, a zero pivot (the
    eps skip) and a non-positive-definite Cholesky."""
    from oracle.oracle import Oracle
    g, o = gpu[prec], Oracle(prec)
    for a in (2.5, -0.75, 0.0):
        A = np.array([[a]])
        L, U, p, s = g.mat_plu(A)
        Lo, Uo, po, so = o.plu(A)
        assert np.array_equal(p, po) and s == so
        assert np.array_equal(L, Lo) and np.array_equal(U, Uo)
        assert g.mat_det(A) == o.det(A)
        Lc, rc = g.mat_chol(A)
        Lco, rco = o.chol(A)
        assert rc == rco
        if rc == 0:
            assert np.allclose(Lc, Lco, rtol=4 * EPS[prec])
        if a != 0.0:
            x = g.mat_solve(A, [1.5])
            assert np.allclose(x, o.solve(A, [1.5]), rtol=4 * EPS[prec])
    C = g.mat_mul([[3.0]], [[-2.0]])
    assert C[0, 0] == -6.0


@pytest.mark.gpu
@pytest.mark.parametrize("prec", PRECS)
def test_qr_against_oracle(gpu, prec):
    """mat_qr / mat_qr_r (mat.h:3907-4192) against the oracle: Q and R to
    8 kappa(A) sqrt(m) eps in the Frobenius norm (both are backward-stable
    Householder QRs with the same sign rule; the reductions differ in
    order), Q orthogonal, A = Q triu(R), the residue below R's diagonal
    within the reference test's m * eps bound; mat_qr_r's R bit for bit
    that of mat_qr.  Shapes: the reference's unblocked sizes (n < 64), one
    and several 32-column blocks with a ragged last block, tall and
    square; one case with a zero column (no reflector)."""
    from oracle.oracle import Oracle
    g, o = gpu[prec], Oracle(prec)
    eps = EPS[prec]
    rng = np.random.default_rng(31)
    for m, n in ((1, 1), (3, 3), (4, 3), (10, 7), (63, 63), (64, 64), (100, 70), (130, 129),
                 (300, 200), (520, 300), (1030, 512), (3000, 40)):
        A = rng.uniform(-1, 1, (m, n)).astype(DT[prec])
        if n > 40:
            A[:, 37] = 0
        Q, R = g.mat_qr(A)
        Qo, Ro = o.qr(A)
        kappa = np.linalg.cond(A[:, np.abs(A).sum(0) > 0].astype(np.float64))
        tol = 8 * kappa * np.sqrt(m) * eps
        assert rel_fro(R, Ro) <= tol, (m, n, rel_fro(R, Ro), tol)
        assert rel_fro(Q, Qo) <= tol, (m, n, rel_fro(Q, Qo), tol)
        assert np.abs(Q.T.astype(np.float64) @ Q - np.eye(m)).max() <= 20 * m * eps, (m, n)
        recon = Q.astype(np.float64) @ np.triu(R).astype(np.float64)
        assert np.abs(recon - A).max() <= 20 * m * eps, (m, n)
        assert np.abs(np.tril(R, -1)).max() <= m * (1e-6 if prec == "f32" else 1e-9), (m, n)
        assert np.array_equal(g.mat_qr_r(A), R), (m, n)
