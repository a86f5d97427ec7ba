"""Shared helpers for the tests: fixtures, device buffers, error metrics."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = {"f32": np.finfo(np.float32).eps, "f64": np.finfo(np.float64).eps}
DT = {"f32": np.float32, "f64": np.float64}

_GOLD = {}


def golden(prec):
    """The reference's recorded inputs and outputs (tests/golden/golden_*.lmfx,
    written by tests/golden/gen_golden.py)."""
    if prec not in _GOLD:
        import sys
        sys.path.insert(0, os.path.join(HERE, "golden"))
        import lmfx
        _GOLD[prec] = lmfx.read(os.path.join(HERE, "golden", f"golden_{prec}.lmfx"))
    return _GOLD[prec]


def rel_fro(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def apply_row_perm(A, p):
    """(PA)[i, :] = A[p[i], :] -- tests/test_mat_plu.c:7-15."""
    return np.asarray(A)[np.asarray(p, dtype=np.int64), :]


def backward_error(A, x, b):
    """Normwise ||Ax - b|| / (||A|| ||x|| + ||b||) in infinity norms, fp64."""
    A = np.asarray(A, np.float64)
    x = np.asarray(x, np.float64).reshape(A.shape[1], -1)
    b = np.asarray(b, np.float64).reshape(A.shape[0], -1)
    r = A @ x - b
    nA = np.abs(A).sum(axis=1).max()
    errs = np.abs(r).max(axis=0) / (nA * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))
    return float(errs.max())


def pivots_match(p_gpu, p_ref, L_ref, prec, c=64.0):
    """Pivot vectors identical except at a near-tie.

    At the first position k where they differ, the GPU's row must have been a
    near-tie candidate in the reference's working column: its multiplier
    |L_ref[i,k]| (i = its final position in the reference) must be within
    c*(k+1)*eps of 1.  Returns (ok, k_first_mismatch or None)."""
    p_gpu = np.asarray(p_gpu)
    p_ref = np.asarray(p_ref)
    diff = np.nonzero(p_gpu != p_ref)[0]
    if diff.size == 0:
        return True, None
    k = int(diff[0])
    pos = {int(r): i for i, r in enumerate(p_ref)}
    i = pos[int(p_gpu[k])]
    tie = abs(1.0 - abs(float(L_ref[i, k]))) <= c * (k + 1) * EPS[prec]
    return tie, k


def to_dev(a, dtype):
    """numpy (rows, cols) -> 1-D cuda tensor holding the column-major data."""
    import torch
    a = np.asarray(a, dtype=dtype)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).cuda()


def from_dev(t, rows, cols):
    return t.cpu().numpy().reshape((rows, cols), order="F")


def stream():
    import torch
    return torch.cuda.current_stream().cuda_stream
