"""Writes the golden fixtures for the hot path from the reference itself.

Runs only where the reference sources exist: it builds oracle/_ref (the
reference's mat.h compiled by oracle/Makefile, x86 scalar path) and records
inputs and outputs of mat_mul / mat_gemm, mat_plu, mat_lu, mat_chol,
mat_solve, mat_solve_spd, the four triangular solves and mat_det, in float
and double builds, as LMFX files next to this script (golden_f32.lmfx,
golden_f64.lmfx; the format is described in lmfx.py and read from C by
lmfx.h).  Inputs are
seeded numpy draws (uniform[-1,1]; SPD inputs G G^T / n + I), plus the
singular and non-positive-definite cases that exercise the reference's
eps-skip (mat.h:4234) and -1 paths (mat.h:6043).

    python tests/golden/gen_golden.py [output directory]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as orc  # noqa: E402

sys.path.insert(0, HERE)
import lmfx  # noqa: E402

PLU_SIZES = [1, 2, 3, 23, 24, 25, 47, 48, 49, 64, 65, 100]
LU_SIZES = [3, 4, 10, 24, 25, 50]
CHOL_SIZES = [1, 2, 31, 32, 33, 64, 65, 100]
TRI_SIZES = [1, 7, 64, 65, 100]
GEMM_SHAPES = [(17, 33, 9), (64, 64, 64), (100, 1, 50), (3, 130, 5)]


def spd(rng, n, dt):
    g = rng.standard_normal((n, n))
    return (g @ g.T / n + np.eye(n)).astype(dt)


def main(outdir=HERE):
    orc.build(ref=True)
    for prec in ("f32", "f64"):
        r = orc.Oracle(prec, kind="ref")
        dt = r.dtype
        rng = np.random.default_rng(20260 + (prec == "f64"))
        out = {}
        if prec == "f32":
            # the CPU-runnable workload: 256x256 mul, plu, solve
            A = rng.uniform(-1, 1, (256, 256)).astype(dt)
            B = rng.uniform(-1, 1, (256, 256)).astype(dt)
            b = rng.uniform(-1, 1, 256).astype(dt)
            L, U, p, s = r.plu(A)
            out.update({"cfg1_A": A, "cfg1_B": B, "cfg1_C": r.mul(A, B), "cfg1_L": L,
                        "cfg1_U": U, "cfg1_p": p, "cfg1_swaps": np.int64(s), "cfg1_b": b,
                        "cfg1_x": r.solve(A, b)})
        for (m, k, n) in GEMM_SHAPES:
            A = rng.uniform(-1, 1, (m, k)).astype(dt)
            B = rng.uniform(-1, 1, (k, n)).astype(dt)
            Cm = rng.uniform(-1, 1, (m, n)).astype(dt)
            tag = f"gemm_{m}_{k}_{n}"
            out.update({f"{tag}_A": A, f"{tag}_B": B, f"{tag}_C0": Cm,
                        f"{tag}_C": r.gemm(0.7, A, B, 0.3, Cm)})
        for n in PLU_SIZES:
            A = rng.uniform(-1, 1, (n, n)).astype(dt)
            b = rng.uniform(-1, 1, n).astype(dt)
            L, U, p, s = r.plu(A)
            out.update({f"plu_{n}_A": A, f"plu_{n}_L": L, f"plu_{n}_U": U, f"plu_{n}_p": p,
                        f"plu_{n}_swaps": np.int64(s), f"plu_{n}_b": b,
                        f"plu_{n}_x": r.solve(A, b), f"plu_{n}_det": np.float64(r.det(A))})
        # singular inputs: near-zero pivots are skipped, not divided by
        sing = {
            "sing3": np.array([[1, 2, 3], [4, 5, 6], [5, 7, 9]], dt),  # tests/test_mat_det.c:55-71
            "zerocol30": None,
            "rank2_40": None,
        }
        z = rng.uniform(-1, 1, (30, 30)).astype(dt)
        z[:, 7] = 0
        sing["zerocol30"] = z
        u = rng.uniform(-1, 1, (40, 2))
        v = rng.uniform(-1, 1, (2, 40))
        sing["rank2_40"] = (u @ v).astype(dt)
        for tag, A in sing.items():
            L, U, p, s = r.plu(A)
            out.update({f"{tag}_A": A, f"{tag}_L": L, f"{tag}_U": U, f"{tag}_p": p,
                        f"{tag}_swaps": np.int64(s), f"{tag}_det": np.float64(r.det(A))})
        for n in LU_SIZES:
            A = rng.uniform(-10, 10, (n, n)).astype(dt)
            L, U, p, q, s = r.lu(A)
            out.update({f"lu_{n}_A": A, f"lu_{n}_L": L, f"lu_{n}_U": U, f"lu_{n}_p": p,
                        f"lu_{n}_q": q, f"lu_{n}_swaps": np.int64(s)})
        for n in CHOL_SIZES:
            A = spd(rng, n, dt)
            b = rng.uniform(-1, 1, n).astype(dt)
            L, rc = r.chol(A)
            x, rs = r.solve_spd(A, b)
            out.update({f"chol_{n}_A": A, f"chol_{n}_L": L, f"chol_{n}_rc": np.int64(rc),
                        f"chol_{n}_b": b, f"chol_{n}_x": x, f"chol_{n}_rs": np.int64(rs)})
        notpd = {"notpd2": np.array([[1, 2], [2, 1]], dt)}  # tests/test_mat_chol.c:133-151
        g = rng.standard_normal((40, 40))
        notpd["indef40"] = ((g + g.T) / 2).astype(dt)
        for tag, A in notpd.items():
            L, rc = r.chol(A)
            x, rs = r.solve_spd(A, np.ones(A.shape[0], dt))
            out.update({f"{tag}_A": A, f"{tag}_rc": np.int64(rc), f"{tag}_rs": np.int64(rs)})
        for n in TRI_SIZES:
            T = rng.uniform(-1, 1, (n, n)).astype(dt) + 4 * np.eye(n, dtype=dt)
            b = rng.uniform(-1, 1, n).astype(dt)
            out.update({f"tri_{n}_T": T, f"tri_{n}_b": b,
                        f"tri_{n}_tril": r.solve_tril(np.tril(T), b),
                        f"tri_{n}_tril_unit": r.solve_tril_unit(np.tril(T), b),
                        f"tri_{n}_triu": r.solve_triu(np.triu(T), b),
                        f"tri_{n}_trilt": r.solve_trilt(np.tril(T), b)})
        path = os.path.join(outdir, f"golden_{prec}.lmfx")
        lmfx.write(path, out)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main(*sys.argv[1:])
