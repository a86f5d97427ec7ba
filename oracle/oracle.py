"""ctypes front end for the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, never by the shipped library.

``Oracle("f32")`` / ``Oracle("f64")`` load the clean-room restatement of
libmat's x86 scalar path (oracle/oracle.c); ``Oracle(..., kind="ref")``
loads the reference itself compiled from its own mat.h into oracle/_ref/
(present only where the reference sources were available at build time).
Both expose the same methods, taking and returning numpy arrays in natural
(row, col) indexing; the column-major marshalling matches libmat's storage
(mat.h:387-393).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def build(ref: bool | None = None) -> None:
    """Compile the restatement, and oracle/_ref when the reference exists."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    refdir = os.environ.get("LIBMAT_REFDIR", "/root/reference")
    if ref is None:
        ref = os.path.exists(os.path.join(refdir, "mat.h"))
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REFDIR={refdir}"], check=True)


def lib_path(prec: str, kind: str = "port") -> str:
    if kind == "ref":
        return os.path.join(HERE, "_ref", f"libmatref_{prec}.so")
    return os.path.join(HERE, "lib", f"liboracle_{prec}.so")


def available(prec: str = "f32", kind: str = "port") -> bool:
    return os.path.exists(lib_path(prec, kind))


class Oracle:
    def __init__(self, prec: str = "f32", kind: str = "port"):
        assert prec in ("f32", "f64") and kind in ("port", "ref")
        path = lib_path(prec, kind)
        if not os.path.exists(path):
            if kind == "port":
                build(ref=False)
            else:
                raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.prec, self.kind = prec, kind
        self.dtype = np.float32 if prec == "f32" else np.float64
        self.ct = C.c_float if prec == "f32" else C.c_double
        tag = ("s" if prec == "f32" else "d")
        self.pre = ("orc_" if kind == "port" else "ref_") + tag + "_"
        f = self._fn
        sz, P, I = C.c_size_t, C.c_void_p, C.c_int
        f("gemm", None, [sz, sz, sz, self.ct, P, P, self.ct, P])
        f("syrk", None, [sz, sz, self.ct, P, self.ct, P, I, I])
        f("plu", I, [sz, P, P, P, P])
        f("lu", I, [sz, P, P, P, P, P])
        f("chol", I, [sz, P, P])
        for t in ("solve_tril", "solve_tril_unit", "solve_triu", "solve_trilt", "solve"):
            f(t, None, [sz, P, P, P])
        f("solve_spd", I, [sz, P, P, P])
        f("solve_multi", None, [sz, sz, P, P, P])
        f("det", self.ct, [sz, P])
        f("inv", None, [sz, P, P])
        f("chol_solve_multi", I, [sz, sz, P, P, P])
        f("batch_solve", None, [sz, P, P, P, P, P, P, P, P])
        f("qr", None, [sz, sz, P, P, P])
        f("qr_r", None, [sz, sz, P, P])

    def _fn(self, name, res, args):
        fn = getattr(self.lib, self.pre + name)
        fn.restype, fn.argtypes = res, args
        setattr(self, "_" + name, fn)

    # -- marshalling: numpy (row, col) array <-> column-major buffer ------
    def _cm(self, a):
        return np.asfortranarray(np.asarray(a, dtype=self.dtype))

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(C.c_void_p)

    def gemm(self, alpha, A, B, beta, Cm):
        A, B = self._cm(A), self._cm(B)
        out = self._cm(Cm).copy(order="F")
        M, K = A.shape
        N = B.shape[1]
        self._gemm(M, K, N, alpha, self._p(A), self._p(B), beta, self._p(out))
        return out

    def syrk(self, alpha, A, beta, Cm, uplo="L", trans=False):
        """C = alpha A A^T + beta C (trans: alpha A^T A + beta C) on the
        uplo triangle; the other triangle of Cm is returned unchanged."""
        A = self._cm(A)
        out = self._cm(Cm).copy(order="F")
        n = A.shape[1] if trans else A.shape[0]
        k = A.shape[0] if trans else A.shape[1]
        assert out.shape == (n, n)
        self._syrk(n, k, alpha, self._p(A), beta, self._p(out), int(uplo in ("U", "u")),
                   int(bool(trans)))
        return out

    def qr(self, A):
        """(Q, R) of the m x n A, m >= n (mat_qr); R keeps the reflectors'
        rounding residue below its diagonal, as libmat leaves it."""
        A = self._cm(A)
        m, n = A.shape
        assert m >= n
        Q = np.zeros((m, m), self.dtype, order="F")
        R = np.zeros((m, n), self.dtype, order="F")
        self._qr(m, n, self._p(A), self._p(Q), self._p(R))
        return Q, R

    def qr_r(self, A):
        """R of the m x n A, m >= n (mat_qr_r)."""
        A = self._cm(A)
        m, n = A.shape
        assert m >= n
        R = np.zeros((m, n), self.dtype, order="F")
        self._qr_r(m, n, self._p(A), self._p(R))
        return R

    def mul(self, A, B):
        A = self._cm(A)
        out = np.zeros((A.shape[0], np.shape(B)[1]), dtype=self.dtype, order="F")
        return self.gemm(1.0, A, B, 0.0, out)

    def plu(self, A):
        A = self._cm(A)
        n = A.shape[0]
        L = np.zeros((n, n), self.dtype, order="F")
        U = np.zeros((n, n), self.dtype, order="F")
        p = np.zeros(n, np.uintp)
        s = self._plu(n, self._p(A), self._p(L), self._p(U), self._p(p))
        return L, U, p.astype(np.int64), s

    def lu(self, A):
        A = self._cm(A)
        n = A.shape[0]
        L = np.zeros((n, n), self.dtype, order="F")
        U = np.zeros((n, n), self.dtype, order="F")
        p = np.zeros(n, np.uintp)
        q = np.zeros(n, np.uintp)
        s = self._lu(n, self._p(A), self._p(L), self._p(U), self._p(p), self._p(q))
        return L, U, p.astype(np.int64), q.astype(np.int64), s

    def chol(self, A):
        A = self._cm(A)
        n = A.shape[0]
        L = np.zeros((n, n), self.dtype, order="F")
        r = self._chol(n, self._p(A), self._p(L))
        return L, r

    def _tri(self, fn, T, b):
        T = self._cm(T)
        b = np.ascontiguousarray(np.asarray(b, self.dtype).reshape(-1))
        x = np.zeros_like(b)
        fn(T.shape[0], self._p(T), self._p(b), self._p(x))
        return x

    def solve_tril(self, L, b):
        return self._tri(self._solve_tril, L, b)

    def solve_tril_unit(self, L, b):
        return self._tri(self._solve_tril_unit, L, b)

    def solve_triu(self, U, b):
        return self._tri(self._solve_triu, U, b)

    def solve_trilt(self, L, b):
        return self._tri(self._solve_trilt, L, b)

    def solve(self, A, b):
        return self._tri(self._solve, A, b)

    def solve_spd(self, A, b):
        A = self._cm(A)
        b = np.ascontiguousarray(np.asarray(b, self.dtype).reshape(-1))
        x = np.zeros_like(b)
        r = self._solve_spd(A.shape[0], self._p(A), self._p(b), self._p(x))
        return x, r

    def solve_multi(self, A, B):
        A, B = self._cm(A), self._cm(B)
        X = np.zeros_like(B, order="F")
        self._solve_multi(A.shape[0], B.shape[1], self._p(A), self._p(B), self._p(X))
        return X

    def inv(self, A):
        A = self._cm(A)
        out = np.zeros_like(A, order="F")
        self._inv(A.shape[0], self._p(A), self._p(out))
        return out

    def chol_solve_multi(self, A, B):
        """mat_chol once + per column mat_solve_tril, mat_solve_trilt;
        returns (X, mat_chol's result)."""
        A, B = self._cm(A), self._cm(B)
        X = np.zeros_like(B, order="F")
        r = self._chol_solve_multi(A.shape[0], B.shape[1], self._p(A), self._p(B), self._p(X))
        return X, r

    def batch_solve(self, n, kind, A, a_off, b, v_off):
        """mat_solve / mat_solve_spd per matrix over a packed batch (column-
        major matrices at A[a_off[i]:], vectors at b[v_off[i]:]); returns
        (x, info)."""
        n = np.ascontiguousarray(n, np.int32)
        kind = np.ascontiguousarray(kind, np.int8)
        A = np.ascontiguousarray(A, self.dtype)
        b = np.ascontiguousarray(b, self.dtype)
        a_off = np.ascontiguousarray(a_off, np.int64)
        v_off = np.ascontiguousarray(v_off, np.int64)
        x = np.zeros_like(b)
        info = np.zeros(n.size, np.int32)
        self._batch_solve(n.size, self._p(n), self._p(kind), self._p(A), self._p(a_off),
                          self._p(b), self._p(x), self._p(v_off), self._p(info))
        return x, info

    def det(self, A):
        A = self._cm(A)
        return float(self._det(A.shape[0], self._p(A)))
