"""Benchmarks of the dense factor-and-solve path on one MI355X.

With no --workload, one run measures every single-GPU configuration and
prints ONE JSON line: the headline (mixed-precision general solve, below)
at the top level and the others under "configs" (fp32 GEMM 8192, fp32
Cholesky 8192 + 8 RHS, mixed-precision SPD 32768, batched 200k), each with
its value, ms_per_step, roofline and cpu_baseline.  Every timed loop runs
with the library's per-launch timing off; the kernel table and the roofline
come from a separate instrumented pass of the same launches afterwards.

Headline workload: mixed-precision general solve.
N = 16384 dense system, 16 right-hand sides ~ U[-1,1], A = U diag(sigma) V^T
with U, V the Q factors of seeded Gaussian matrices and sigma geometrically
spaced for condition numbers kappa in {1e2, 1e4, 1e6}.  One step = solving
all three systems with mat_gpu_gesv_mixed (fp32 working LU with split-BF16
MFMA trailing updates + fp64 iterative refinement to backward error
<= 1e-14).  The comparison is the plain fp64 GPU LU solve (mat_gpu_gesv_f64)
on the same systems, plus torch.linalg.solve (the vendor fp64 solver) for
reference.  Both are credited with the same nominal FLOPs,
2/3 n^3 + 2 n^2 * nrhs per system (the reference's own FLOP model,
tests/bench/compare/bench_solve.cpp:79).  Inputs are resident in HBM when
the timed region starts.

Other workloads (--workload):
  gemm       fp32 C = A B, 8192^3 (mat_gemm's GPU form)
  chol       fp32 Cholesky solve, N = 8192, 8 right-hand sides, SPD
             A = G G^T / N + I
  mixed-spd  N = 32768 SPD, A = Q diag(lambda) Q^T, lambda log-uniform on
             [1, kappa], kappa in {1e2, 1e3}, 4 right-hand sides ~ N(0,1):
             mat_gpu_posv_mixed (FP8 tile-scaled trailing updates + fp64
             refinement) against mat_gpu_posv_f64
  batched    200000 systems with n log-uniform on [3, 64], half general
             (U[-1,1] + n I), half SPD (G G^T + I), one right-hand side
             each, fp32: one persistent kernel per size bin and kind,
             captured in one hipGraph and replayed, against one launch per
             matrix

Every line carries "roofline" for the workload's dominant kernel family:
achieved = algorithmic bytes (or flops) of its launches / their summed
durations, measured with HIP events recorded by the library on the launch
stream during an instrumented pass after the timed steps, with "traffic" =
HBM bytes per launch from
the rocprofv3 FETCH_SIZE / WRITE_SIZE passes committed under profiles/
(hbm_traffic__*.md; null for a workload run at other sizes or not
profiled); and "cpu_baseline": the reference's x86
scalar path (oracle/_ref, compiled from the reference sources; the
clean-room oracle restatement when that is absent) timed in a child process
pinned to one host core while the GPU runs: 1 s warm-up, then the median of
10 samples at three sizes, fitted as a n^3 and reported as extrapolated to
the configuration's size (the batched loop runs at full size).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published dense peaks
PEAK = {"bf16": 2500.0, "fp8": 5000.0, "fp32": 157.3, "fp64": 78.6, "hbm": 8000.0}
FAMILIES = ["panel", "rowswap", "gemm", "gemm_bf16x", "gemm_fp8", "pack", "trsm", "tri_chain",
            "chol_leaf", "chol_panel", "qr_panel"]
# family -> (bound, unit, peak key for the given dtype)
BOUND = {"panel": ("hbm", "GB/s"), "rowswap": ("hbm", "GB/s"), "pack": ("hbm", "GB/s"),
         "tri_chain": ("hbm", "GB/s"), "chol_leaf": ("hbm", "GB/s"), "trsm": ("hbm", "GB/s"),
         "chol_panel": ("hbm", "GB/s"),
         "qr_panel": ("hbm", "GB/s"),
         "gemm": ("mfma", "TFLOP/s"), "gemm_bf16x": ("mfma", "TFLOP/s"),
         "gemm_fp8": ("mfma", "TFLOP/s")}


def ptr(t):
    return C.c_void_p(t.data_ptr())


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------ kernel timing
def kernel_table(lib):
    out = {}
    for fam in FAMILIES:
        n = C.c_int64(0)
        ms = C.c_double(0.0)
        work = C.c_double(0.0)
        rc = lib.mat_gpu_timing_query(fam.encode(), C.byref(n), C.byref(ms), C.byref(work))
        assert rc == 0
        if n.value:
            out[fam] = {"launches": n.value, "ms": round(ms.value, 3), "work": work.value}
    return out


def hbm_traffic(family, workload):
    """HBM bytes per launch of one kernel family on one workload, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes committed under profiles/
    (hbm_traffic__<kernels>__<workload>.md, calibrated against a copy kernel
    of known size).  None when no pass covers it or the workload was run at
    other sizes than the profiled one."""
    if workload is None:
        return None
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"hbm_traffic__*__{workload}.md"))):
        with open(path) as f:
            rows = [ln.strip().strip("|").split("|") for ln in f if ln.startswith("|")]
        if not rows:
            continue
        head = [c.strip() for c in rows[0]]
        if "family" not in head or "HBM bytes per launch" not in head:
            continue
        fi, ti = head.index("family"), head.index("HBM bytes per launch")
        for r in rows[1:]:
            if len(r) > max(fi, ti) and r[fi].strip() == family:
                try:
                    return int(float(r[ti].strip()))
                except ValueError:
                    return None
    return None


def roofline(table, family, gemm_dtype="fp32", steps=1, workload=None):
    """Roofline object for one kernel family from the library's HIP-event table."""
    rec = table[family]
    bound, unit = BOUND[family]
    sec = rec["ms"] * 1e-3
    if bound == "hbm":
        achieved = rec["work"] / sec / 1e9
        peak = PEAK["hbm"]
    else:
        achieved = rec["work"] / sec / 1e12
        peak = PEAK[{"gemm_bf16x": "bf16", "gemm_fp8": "fp8"}.get(family, gemm_dtype)]
    return {"kernel": family, "bound": bound, "achieved": round(achieved, 3), "peak": peak,
            "unit": unit, "frac": round(achieved / peak, 5),
            "traffic": hbm_traffic(family, workload),
            "launches": rec["launches"], "avg_launch_us": round(rec["ms"] * 1e3 / rec["launches"], 3),
            "ms_per_step": round(rec["ms"] / steps, 3)}


def dominant(table, exclude=()):
    fams = [f for f in table if f not in exclude and table[f]["work"] > 0]
    return max(fams, key=lambda f: table[f]["ms"])


def breakdown(table, steps):
    return {f: {"launches": r["launches"] // steps, "ms_per_step": round(r["ms"] / steps, 3)}
            for f, r in sorted(table.items(), key=lambda x: -x[1]["ms"])}


def instrumented_pass(torch, g, step, reps):
    """Run `step` reps times with the library's per-launch HIP-event timing
    on (after the timed loop, which runs with it off) and return the kernel
    table of those launches."""
    torch.cuda.synchronize()
    g.check(g.lib.mat_gpu_timing_enable(1), "timing")
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    g.check(g.lib.mat_gpu_timing_enable(0), "timing")
    return kernel_table(g.lib)


def timed(torch, step, steps, warmup, label):
    """Warm-up, then `steps` steps bracketed by device synchronizations;
    returns the wall seconds of the timed steps."""
    for w in range(warmup):
        step()
        log(f"{label}: warmup {w} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step()
        if steps <= 25:
            log(f"{label}: step {k} done")
    torch.cuda.synchronize()
    return time.perf_counter() - t0


INSTR_NOTE = ("kernel durations from HIP events the library records on each launch's stream, "
              "in an instrumented pass of {} step(s) after the timed loop")


# ------------------------------------------------------------ CPU baselines
# Run in a child process (`bench.py --cpu-worker`) pinned to one host core,
# started before this process touches the GPU and timed while the GPU runs.
# Method of the reference's benchmark harness (zap.h:792-804, 1133-1226):
# monotonic clock, 1 s warm-up, median of 10 samples.  The scalar path is
# far too slow for the configurations' sizes (GEMM 8192 alone would take
# over half an hour), so each op is sampled at three sizes, fitted as
# t = a n^3 and reported as extrapolated to the configuration's size; the
# batched per-matrix loop runs at full size.
CPU_SAMPLES = 10


def _host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def _oracle(prec):
    from oracle import oracle as orc
    kind = "reference" if orc.available(prec, "ref") else "port"
    return orc.Oracle(prec, kind="ref" if kind == "reference" else "port"), kind


def _median_s(fn, samples):
    ts = []
    for _ in range(samples):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def _warm_up(fn, secs=1.0):
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < secs:
        fn()


def _cpu_sweep(make, flops, sizes, config_n, what):
    """make(n) -> a no-argument callable doing one sample's work at order n."""
    _warm_up(make(sizes[0]))
    points = []
    for n in sizes:
        fn = make(n)
        fn()
        t = _median_s(fn, CPU_SAMPLES)
        points.append({"n": n, "median_s": round(t, 6), "gflops": round(flops(n) / t / 1e9, 4),
                       "samples": CPU_SAMPLES})
    ln_n = np.log(np.array(sizes, dtype=np.float64))
    ln_t = np.log(np.array([p["median_s"] for p in points]))
    a = float(np.exp(np.mean(ln_t - 3.0 * ln_n)))
    t_cfg = a * config_n ** 3
    return {"value": round(flops(config_n) / t_cfg / 1e9, 4), "unit": "GFLOP/s",
            "extrapolated": True, "config_n": config_n,
            "extrapolated_s_at_config": round(t_cfg, 2),
            "fit": "t = a n^3 over the sampled sizes",
            "fit_free_exponent": round(float(np.polyfit(ln_n, ln_t, 1)[0]), 3),
            "points": points, "sample": what}


def cpu_op(name):
    rng = np.random.default_rng(0)
    if name == "mixed-lu":
        o, kind = _oracle("f64")
        nrhs = 16

        def make(n):
            A, B = rng.uniform(-1, 1, (n, n)), rng.uniform(-1, 1, (n, nrhs))
            return lambda: o.solve_multi(A, B)
        r = _cpu_sweep(make, lambda n: 2.0 / 3.0 * n ** 3 + 2.0 * n ** 2 * nrhs, [512, 1024, 2048],
                       16384, "mat_plu + 16 x (permute, mat_solve_tril_unit, mat_solve_triu), "
                              "fp64, U[-1,1] systems (a direct LU's time does not depend on kappa)")
    elif name == "gemm":
        o, kind = _oracle("f32")

        def make(n):
            A = rng.uniform(-1, 1, (n, n)).astype(np.float32)
            B = rng.uniform(-1, 1, (n, n)).astype(np.float32)
            return lambda: o.mul(A, B)
        r = _cpu_sweep(make, lambda n: 2.0 * n ** 3, [256, 384, 512], 8192,
                       "mat_mul fp32, U[-1,1]")
    elif name in ("chol", "mixed-spd"):
        prec, nrhs, cfg = ("f32", 8, 8192) if name == "chol" else ("f64", 4, 32768)
        o, kind = _oracle(prec)

        def make(n):
            G = rng.standard_normal((n, n))
            A = (G @ G.T / n + np.eye(n)).astype(o.dtype)
            B = rng.uniform(-1, 1, (n, nrhs)).astype(o.dtype)
            return lambda: o.chol_solve_multi(A, B)
        r = _cpu_sweep(make, lambda n: n ** 3 / 3.0 + 2.0 * n ** 2 * nrhs, [512, 768, 1024], cfg,
                       f"mat_chol + {nrhs} x (mat_solve_tril, mat_solve_trilt), {prec}, "
                       "SPD G G^T / n + I")
    elif name == "batched":
        o, kind = _oracle("f32")
        ns, kinds, A, b, aoff, voff = make_batch(200000)
        _warm_up(lambda: o.batch_solve(ns[:2000], kinds[:2000], A, aoff[:2000], b, voff[:2000]))
        t = _median_s(lambda: o.batch_solve(ns, kinds, A, aoff, b, voff), CPU_SAMPLES)
        flops = batch_flops(ns, kinds)
        r = {"value": round(flops / t / 1e9, 4), "unit": "GFLOP/s",
             "matrices_per_s": round(ns.size / t, 1), "extrapolated": False,
             "median_s": round(t, 4), "samples": CPU_SAMPLES,
             "sample": "all 200000 systems, one mat_solve / mat_solve_spd per matrix in a C loop"}
    elif name == "qr":
        o, kind = _oracle("f32")

        def make(n):
            A = rng.uniform(-1, 1, (n, n)).astype(np.float32)
            return lambda: o.qr(A)
        r = _cpu_sweep(make, lambda n: 8.0 * n ** 3 / 3.0, [256, 384, 512], 4096,
                       "mat_qr (R and Q), fp32, U[-1,1]")
    else:
        raise ValueError(name)
    r["kind"] = kind
    return r


def cpu_worker(ops, core):
    """Child-process entry: pin to one core, time the ops, print one JSON
    object {op: baseline} on stdout.  Never touches the GPU."""
    if core is not None:
        os.sched_setaffinity(0, {core})
    host = {"cpu_model": _host_cpu(), "nproc": os.cpu_count(), "cores_used": 1,
            "pinned_core": core}
    out = {}
    for op in ops:
        t0 = time.perf_counter()
        try:
            out[op] = cpu_op(op)
        except Exception as e:  # reported in the JSON, the GPU results still stand
            out[op] = {"error": repr(e)}
        out[op].update({"cores": 1, "host": host})
        log(f"cpu baseline {op} done in {time.perf_counter() - t0:.1f} s")
    print(json.dumps(out), flush=True)


class CpuBaselines:
    """The CPU-baseline child process, started before the GPU is touched."""

    def __init__(self, ops):
        import subprocess
        self.proc = None
        if not ops:
            return
        cores = sorted(os.sched_getaffinity(0))
        core = cores[-1]
        self.proc = subprocess.Popen(
            [sys.executable, os.path.abspath(__file__), "--cpu-worker", ",".join(ops),
             "--cpu-core", str(core)], stdout=subprocess.PIPE, text=True)

    def results(self, timeout=900):
        if self.proc is None:
            return {}
        out, _ = self.proc.communicate(timeout=timeout)
        try:
            return json.loads(out.strip().splitlines()[-1])
        except (ValueError, IndexError):
            return {"error": f"cpu worker exit {self.proc.returncode}"}


# ------------------------------------------------------------ mixed-lu (default)
OWN_QR = False  # --qr own: orthogonal factors from this library's QR


def q_of_gaussian(torch, n, gen):
    """Q factor of a seeded n x n Gaussian (fp64): the vendor QR, or with
    --qr own this library's mat_gpu_geqrf_f64 (same family of matrices;
    used where the vendor QR does not run, such as under counter profiling)."""
    G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen)
    if not OWN_QR:
        Q, _ = torch.linalg.qr(G)
        return Q
    import libmat_amd
    g = libmat_amd.load("f64")
    Qb = torch.empty_like(G)
    # G's row-major buffer is a column-major matrix too (its transpose):
    # another Gaussian; the column-major buffer of its Q is Q^T row-major
    g.check(g.lib.mat_gpu_geqrf_f64(torch.cuda.current_stream().cuda_stream, n, n, ptr(G), n,
                                    ptr(Qb)), "geqrf")
    torch.cuda.synchronize()
    log("own QR done")
    return Qb.T


def make_general(torch, n, kappas, nrhs, seed=2026):
    """A_kappa = U diag(sigma_kappa) V^T (fp64, column-major buffers), one B."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    U = q_of_gaussian(torch, n, gen)
    V = q_of_gaussian(torch, n, gen)
    mats = []
    for kappa in kappas:
        sig = torch.logspace(0, -np.log10(kappa), n, dtype=torch.float64, device="cuda")
        # the column-major buffer of (U diag(sig) V^T)^T = V diag(sig) U^T is the
        # row-major tensor U diag(sig) V^T: either way a member of the same family
        mats.append(((U * sig) @ V.T).contiguous())
    del U, V
    B = (torch.rand(nrhs, n, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1)
    torch.cuda.synchronize()
    return mats, B.contiguous()


def make_uniform(torch, n, count, nrhs, seed=2028):
    """count systems with entries ~ U[-1,1] (fp64) and one B: the same shapes
    and launches as the conditioned systems without a QR factorization (for
    counter profilers that cannot run the vendor QR)."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    mats = [(torch.rand(n, n, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1)
            for _ in range(count)]
    B = (torch.rand(nrhs, n, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1)
    torch.cuda.synchronize()
    return mats, B.contiguous()


def berr_fp64(torch, A, X, B):
    """Normwise backward error (inf norms) with A as a column-major buffer."""
    n = A.shape[0]
    Am = A.T  # logical column-major matrix
    Xm = X.reshape(-1, n).T
    Bm = B.reshape(-1, n).T
    R = Am @ Xm - Bm
    nA = Am.abs().sum(dim=1).max()
    e = R.abs().amax(dim=0) / (nA * Xm.abs().amax(dim=0) + Bm.abs().amax(dim=0))
    return float(e.max())


PIVOTING = {0: "partial pivoting", 1: "random butterfly, no pivoting",
            2: "tournament pivoting"}


def run_mixed(torch, libmat_amd, args, spd):
    g = libmat_amd.load("f64")
    lib = g.lib
    n, nrhs = args.n, args.nrhs
    kappas = [float(k) for k in args.kappas.split(",")]
    log(f"generating {len(kappas)} {'SPD' if spd else 'general'} systems n={n}")
    if args.data == "uniform":
        mats, B = make_uniform(torch, n, len(kappas), nrhs)
    else:
        mats, B = (make_spd(torch, n, kappas, nrhs) if spd else
                   make_general(torch, n, kappas, nrhs))
    X = torch.zeros(nrhs * n, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    opts = libmat_amd.IrOpts()
    lib.mat_gpu_ir_defaults(C.byref(opts), 1 if spd else 0)
    if args.passes is not None:
        opts.passes = args.passes
    if args.pivoting is not None and not spd:
        opts.pivoting = args.pivoting
    if args.leaf is not None:
        opts.leaf = args.leaf
    status = [libmat_amd.IrStatus() for _ in kappas]
    solve = lib.mat_gpu_posv_mixed if spd else lib.mat_gpu_gesv_mixed
    plain = lib.mat_gpu_posv_f64 if spd else lib.mat_gpu_gesv_f64

    def step():
        for i, A in enumerate(mats):
            g.check(solve(stream, n, nrhs, ptr(A), n, ptr(B), n, ptr(X), n, C.byref(opts),
                          C.byref(status[i])), "mixed solve")

    label = "mixed-spd" if spd else "mixed-lu"
    elapsed = timed(torch, step, args.steps, args.warmup, label)
    table = instrumented_pass(torch, g, step, 1)
    ms_per_step = elapsed / args.steps * 1e3
    if spd:
        flops_per_solve = n ** 3 / 3.0 + 2.0 * n ** 2 * nrhs
    else:
        flops_per_solve = 2.0 / 3.0 * n ** 3 + 2.0 * n ** 2 * nrhs
    value = flops_per_solve * len(kappas) * args.steps / elapsed / 1e9

    # per-kappa accuracy and the fp64 comparisons (outside the timed region)
    per = []
    Xp = torch.zeros_like(X)
    plain_ms_total = 0.0
    vendor_ms_total = 0.0
    for i, (kappa, A) in enumerate(zip(kappas, mats)):
        g.check(solve(stream, n, nrhs, ptr(A), n, ptr(B), n, ptr(X), n, C.byref(opts),
                      C.byref(status[i])), "mixed solve")
        st = status[i]
        entry = {"kappa": kappa, "iterations": st.iterations, "fell_back": bool(st.fell_back),
                 "backward_error": berr_fp64(torch, A, X, B),
                 "lib_backward_error": st.backward_error, "factor_ms": round(st.factor_ms, 3),
                 "solve_ms": round(st.total_ms, 3)}
        if not args.no_plain:
            torch.cuda.synchronize()
            tp = time.perf_counter()
            g.check(plain(stream, n, nrhs, ptr(A), n, ptr(B), n, ptr(Xp), n), "plain fp64")
            torch.cuda.synchronize()
            tp = time.perf_counter() - tp
            plain_ms_total += tp * 1e3
            Xm, Xr = X.reshape(nrhs, n), Xp.reshape(nrhs, n)
            entry["plain_fp64_ms"] = round(tp * 1e3, 3)
            entry["plain_fp64_backward_error"] = berr_fp64(torch, A, Xp, B)
            entry["forward_error_vs_fp64"] = float((Xm - Xr).abs().max() / Xr.abs().max())
            # vendor fp64 solver through torch (rocSOLVER / hipSOLVER), same system
            Am, Bm = A.T, B.T
            torch.cuda.synchronize()
            tv = time.perf_counter()
            if spd:
                Xv = torch.cholesky_solve(Bm, torch.linalg.cholesky(Am))
            else:
                Xv = torch.linalg.solve(Am, Bm)
            torch.cuda.synchronize()
            tv = time.perf_counter() - tv
            vendor_ms_total += tv * 1e3
            entry["vendor_fp64_ms"] = round(tv * 1e3, 3)
            entry["vendor_fp64_backward_error"] = berr_fp64(torch, A, Xv.T.contiguous(), B)
            del Xv
        per.append(entry)
        log(f"kappa={kappa:g}: {entry}")

    name = "mixed-precision SPD solve" if spd else "mixed-precision general solve"
    out = {"metric": "solve GFLOP/s: mixed-prec LU/Chol+IR vs plain FP64 at equal backward error",
           "value": round(value, 2), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": ("f32+fp8 factorization, f64 refinement" if spd else
                     "f32+bf16 factorization, f64 refinement"),
           "data": ("synthetic (seeded; U[-1,1] entries, kappa not controlled)"
                    if args.data == "uniform" else
                    "synthetic (seeded; " + ("Q diag(lambda) Q^T, Q = QR of Gaussian" if spd else
                                             "U diag(sigma) V^T, U,V = QR of Gaussian")
                    + (" by this library's QR" if OWN_QR else "") + ")"),
           "config": {"workload": name, "n": n, "nrhs": nrhs, "kappas": kappas,
                      "factorization": ("fp32 Cholesky, FP8 e4m3 per-128x128-tile-scaled MFMA "
                                        "trailing updates" if spd else
                                        f"fp32 LU ({PIVOTING[opts.pivoting]}), BF16 x{opts.passes} "
                                        "split MFMA trailing updates"),
                      "refinement": "fp64 residual + fp64 update, fallback to fp64 on stall",
                      "parallelism": "single GPU"},
           "per_kappa": per, "kernels": breakdown(table, 1)}
    if not args.no_plain:
        plain_gflops = flops_per_solve * len(kappas) / (plain_ms_total * 1e-3) / 1e9
        out["plain_fp64"] = {"value": round(plain_gflops, 2), "unit": "GFLOP/s",
                             "ms_per_step": round(plain_ms_total, 3)}
        out["speedup_vs_plain_fp64"] = round(value / plain_gflops, 3)
        vendor_gflops = flops_per_solve * len(kappas) / (vendor_ms_total * 1e-3) / 1e9
        out["vendor_fp64"] = {"value": round(vendor_gflops, 2), "unit": "GFLOP/s",
                              "ms_per_step": round(vendor_ms_total, 3),
                              "solver": ("torch.linalg.cholesky + cholesky_solve (fp64)" if spd
                                         else "torch.linalg.solve (fp64)")}
    del mats
    torch.cuda.empty_cache()
    dom = dominant(table)
    out["roofline"] = roofline(table, dom, steps=1, workload=args.profiled)
    out["roofline"]["measured_in"] = INSTR_NOTE.format(1)
    gem = "gemm_fp8" if spd else "gemm_bf16x"
    if gem in table and gem != dom:
        out["roofline_trailing_update"] = roofline(table, gem, steps=1, workload=args.profiled)
    return out


def orthonormal(torch, n, gen):
    """Q factor (positive-diagonal R) of an n x n Gaussian by CholeskyQR2:
    Q = G R^-1 with R = chol(G^T G), applied twice; each pass is a GEMM, a
    Cholesky and a triangular solve, logged separately."""
    Q = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen)
    for it in range(2):
        R = torch.linalg.cholesky(Q.T @ Q, upper=True)
        Q = torch.linalg.solve_triangular(R, Q, upper=True, left=False)
        del R
        torch.cuda.synchronize()
        log(f"orthonormalization pass {it} done")
    return Q


def make_spd(torch, n, kappas, nrhs, seed=2027):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    Q = q_of_gaussian(torch, n, gen) if OWN_QR else orthonormal(torch, n, gen)
    mats = []
    for kappa in kappas:
        u = torch.rand(n, dtype=torch.float64, device="cuda", generator=gen)
        lam = torch.exp(u * np.log(kappa))
        A = (Q * lam) @ Q.T
        A = (A + A.T) * 0.5
        mats.append(A.contiguous())
        del A
        torch.cuda.synchronize()
        log(f"kappa={kappa:g} system built")
    del Q
    B = torch.randn(nrhs, n, dtype=torch.float64, device="cuda", generator=gen)
    torch.cuda.synchronize()
    return mats, B.contiguous()


# ------------------------------------------------------------ gemm
def run_gemm(torch, libmat_amd, args):
    g = libmat_amd.load("f32")
    lib = g.lib
    n = args.n
    gen = torch.Generator(device="cuda").manual_seed(7)
    A = torch.rand(n * n, device="cuda", generator=gen) * 2 - 1
    B = torch.rand(n * n, device="cuda", generator=gen) * 2 - 1
    Cc = torch.zeros(n * n, device="cuda")
    s = torch.cuda.current_stream().cuda_stream

    def step():
        g.check(lib.mat_gpu_gemm_f32(s, 0, 0, n, n, n, 1.0, ptr(A), n, ptr(B), n, 0.0, ptr(Cc), n),
                "gemm")

    elapsed = timed(torch, step, args.steps, args.warmup, "gemm")
    table = instrumented_pass(torch, g, step, 3)
    # sampled check against an fp64 product of the same operands
    rows = torch.randint(0, n, (64,), device="cuda", generator=gen)
    Am = A.view(n, n).T  # column-major buffers
    Bm = B.view(n, n).T
    ref = Am[rows].double() @ Bm.double()
    got = Cc.view(n, n).T[rows].double()
    rel = float((got - ref).abs().max() / ref.abs().max())
    value = 2.0 * n ** 3 * args.steps / elapsed / 1e9
    out = {"metric": "fp32 GEMM GFLOP/s", "value": round(value, 2), "unit": "GFLOP/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic U[-1,1]",
           "config": {"workload": "fp32 GEMM (mat_gemm)", "m": n, "n": n, "k": n,
                      "parallelism": "single GPU"},
           "max_rel_err_vs_fp64_sampled_rows": rel, "kernels": breakdown(table, 3),
           "roofline": roofline(table, "gemm", "fp32", 3, args.profiled)}
    out["roofline"]["measured_in"] = INSTR_NOTE.format(3)
    return out


# ------------------------------------------------------------ chol
def run_chol(torch, libmat_amd, args):
    g = libmat_amd.load("f32")
    lib = g.lib
    n, nrhs = args.n, args.nrhs
    gen = torch.Generator(device="cuda").manual_seed(11)
    G = torch.randn(n, n, device="cuda", generator=gen)
    A = (G @ G.T / n + torch.eye(n, device="cuda")).contiguous()
    del G
    B = (torch.rand(nrhs, n, device="cuda", generator=gen) * 2 - 1).contiguous()
    X = torch.zeros_like(B)
    s = torch.cuda.current_stream().cuda_stream

    def step():
        rc = lib.mat_gpu_posv_f32(s, n, nrhs, ptr(A), n, ptr(B), n, ptr(X), n)
        g.check(rc, "posv")
        assert rc == 0, "matrix not positive definite"

    elapsed = timed(torch, step, args.steps, args.warmup, "chol")
    table = instrumented_pass(torch, g, step, 3)
    Ad = A.double()
    R = Ad @ X.double().T - B.double().T
    berr = float((R.abs().amax(dim=0) / (Ad.abs().sum(dim=1).max() * X.double().abs().amax(dim=1)
                                         + B.double().abs().amax(dim=1))).max())
    flops = n ** 3 / 3.0 + 2.0 * n ** 2 * nrhs
    value = flops * args.steps / elapsed / 1e9
    out = {"metric": "fp32 Cholesky solve GFLOP/s", "value": round(value, 2), "unit": "GFLOP/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic SPD G G^T / N + I, G ~ N(0,1)",
           "config": {"workload": "fp32 Cholesky solve (mat_chol + mat_solve_spd)", "n": n,
                      "nrhs": nrhs, "parallelism": "single GPU"},
           "backward_error": berr, "kernels": breakdown(table, 3),
           "roofline": roofline(table, dominant(table), "fp32", 3, args.profiled)}
    out["roofline"]["measured_in"] = INSTR_NOTE.format(3)
    return out


# ------------------------------------------------------------ qr
def run_qr(torch, libmat_amd, args):
    """mat_qr (A = Q R, R and the full Q) of an n x n fp32 matrix held in HBM,
    through the device entry point mat_gpu_geqrf_f32 (the kernels mat_qr
    runs).  A step restores A from a resident copy (2 n^2 x 4 bytes), then
    factors it."""
    g = libmat_amd.load("f32")
    lib = g.lib
    n = args.n
    gen = torch.Generator(device="cuda").manual_seed(13)
    A0 = (torch.rand(n, n, device="cuda", generator=gen) * 2 - 1).contiguous()
    A = torch.empty_like(A0)
    Q = torch.empty_like(A0)
    s = torch.cuda.current_stream().cuda_stream

    def step(q=True):
        A.copy_(A0)
        g.check(lib.mat_gpu_geqrf_f32(s, n, n, ptr(A), n, ptr(Q) if q else None), "geqrf")

    elapsed = timed(torch, step, args.steps, args.warmup, "qr")
    table = instrumented_pass(torch, g, step, 1)
    # column-major buffers: the factored matrix is A0^T, Q and R likewise
    M = A0.double().T
    Qm, Rm = Q.double().T, torch.triu(A.double().T)
    recon = float((Qm @ Rm - M).norm() / M.norm())
    orth = float((Qm.T @ Qm - torch.eye(n, dtype=torch.float64, device="cuda")).abs().max())
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step(q=False)
    torch.cuda.synchronize()
    r_ms = (time.perf_counter() - t1) / args.steps * 1e3
    flops = 8.0 * n ** 3 / 3.0
    value = flops * args.steps / elapsed / 1e9
    out = {"metric": "fp32 Householder QR GFLOP/s (mat_qr: R and Q)", "value": round(value, 2),
           "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic U[-1,1] square matrix",
           "config": {"workload": "fp32 Householder QR (mat_qr, full Q)", "n": n,
                      "flops": "8/3 n^3 (R 4/3 n^3 + Q 4/3 n^3)", "parallelism": "single GPU"},
           "reconstruction_error": recon, "orthogonality_error": orth,
           "qr_r_ms": round(r_ms, 3), "qr_r_gflops": round(4.0 * n ** 3 / 3.0 / r_ms / 1e6, 2),
           "kernels": breakdown(table, 1),
           # the panel is the critical path (the caller's stream); the GEMMs,
           # about as long in sum, run mostly on the side stream beside it
           "roofline": roofline(table, "qr_panel", "fp32", 1, args.profiled),
           "roofline_updates": roofline(table, "gemm", "fp32", 1, args.profiled)}
    # one workgroup per column step: latency-bound, reported against HBM
    out["roofline"]["bound"] = "latency (reported against hbm)"
    out["roofline"]["measured_in"] = INSTR_NOTE.format(1)
    return out


# ------------------------------------------------------------ batched
def make_batch(count, seed=5, lo=3, hi=64):
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
            Mf = np.transpose(M, (0, 2, 1)).reshape(idx.size, n * n).astype(np.float32)  # column-major
            for j, i in enumerate(idx):
                A[aoff[i]:aoff[i] + n * n] = Mf[j]
    return ns, kinds, A, b, aoff, voff


def batch_flops(ns, kinds):
    """Sum of (2/3 n^3 + 2 n^2) over the general systems and (n^3/3 + 2 n^2)
    over the SPD ones (bench_solve.cpp:79, bench_solve_spd.cpp:89)."""
    n64 = np.asarray(ns, np.float64)
    return float(np.sum(np.where(np.asarray(kinds) == 0, 2.0 / 3.0 * n64 ** 3 + 2 * n64 ** 2,
                                 n64 ** 3 / 3.0 + 2 * n64 ** 2)))


def run_batched(torch, libmat_amd, args):
    g = libmat_amd.load("f32")
    lib = g.lib
    count = args.count
    log(f"generating {count} systems")
    ns, kinds, A, b, aoff, voff = make_batch(count)
    d = {k: torch.from_numpy(v).cuda() for k, v in (("A", A), ("b", b), ("aoff", aoff),
                                                      ("voff", voff))}
    x = torch.zeros(b.size, dtype=torch.float32, device="cuda")
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    plan = C.c_void_p()
    g.check(lib.mat_gpu_batch_create_f32(s, count, ns.ctypes.data_as(C.c_void_p),
                                         kinds.ctypes.data_as(C.c_void_p), ptr(d["A"]),
                                         ptr(d["aoff"]), ptr(d["b"]), ptr(x), ptr(d["voff"]),
                                         ptr(info), 1, C.byref(plan)), "batch plan")
    wg, order, tuned = C.c_int(), C.c_int(), C.c_double()
    g.check(lib.mat_gpu_batch_tuning(plan, C.byref(wg), C.byref(order), C.byref(tuned)), "tuning")
    bin_mask, untuned = C.c_int(), C.c_double()
    waves = (C.c_int32 * 8)()
    g.check(lib.mat_gpu_batch_tuning_shape(plan, C.byref(bin_mask), waves, C.byref(untuned)),
            "tuning shape")

    def step():
        g.check(lib.mat_gpu_batch_run(s, plan), "batch run")

    for w in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    batch_ms = e0.elapsed_time(e1) / args.steps
    assert int((info != 0).sum()) == 0
    # residual check on a sample
    xh = x.cpu().numpy()
    worst = 0.0
    for i in range(0, count, max(1, count // 500)):
        n = ns[i]
        M = A[aoff[i]:aoff[i] + n * n].reshape(n, n, order="F").astype(np.float64)
        xi = xh[voff[i]:voff[i] + n].astype(np.float64)
        bi = b[voff[i]:voff[i] + n].astype(np.float64)
        r = np.abs(M @ xi - bi).max() / (np.abs(M).sum(1).max() * np.abs(xi).max() + np.abs(bi).max())
        worst = max(worst, r)

    # per-matrix baselines on a sample: one plan (a one-matrix persistent launch,
    # replayed from its graph) per matrix, and the single-system entry points
    sample = min(args.loop_sample, count)
    plans = []
    for i in range(sample):
        p = C.c_void_p()
        nn = ns[i:i + 1].copy()
        kk = kinds[i:i + 1].copy()
        ao = torch.tensor([0], dtype=torch.int64, device="cuda")
        sub_a = d["A"][aoff[i]:aoff[i] + nn[0] ** 2]
        sub_b = d["b"][voff[i]:voff[i] + nn[0]]
        sub_x = x[voff[i]:voff[i] + nn[0]]
        g.check(lib.mat_gpu_batch_create_f32(s, 1, nn.ctypes.data_as(C.c_void_p),
                                             kk.ctypes.data_as(C.c_void_p), ptr(sub_a), ptr(ao),
                                             ptr(sub_b), ptr(sub_x), ptr(ao), ptr(info[i:i + 1]), 0,
                                             C.byref(p)), "plan")
        plans.append((p, ao))
    for p, _ in plans[:10]:
        lib.mat_gpu_batch_run(s, p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p, _ in plans:
        lib.mat_gpu_batch_run(s, p)
    torch.cuda.synchronize()
    per_launch_s = (time.perf_counter() - t0) / sample
    for p, _ in plans:
        lib.mat_gpu_batch_destroy(p)
    sample2 = min(200, count)
    xs = torch.zeros(64, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(sample2):
        n = int(ns[i])
        sa = d["A"][aoff[i]:aoff[i] + n * n]
        sb = d["b"][voff[i]:voff[i] + n]
        if kinds[i] == 0:
            lib.mat_gpu_gesv_f32(s, n, 1, ptr(sa), n, ptr(sb), n, ptr(xs), n)
        else:
            lib.mat_gpu_posv_f32(s, n, 1, ptr(sa), n, ptr(sb), n, ptr(xs), n)
    torch.cuda.synchronize()
    per_call_s = (time.perf_counter() - t0) / sample2
    lib.mat_gpu_batch_destroy(plan)

    n64 = ns.astype(np.float64)
    flops = batch_flops(ns, kinds)
    bytes_min = float(np.sum(n64 ** 2 + 2 * n64) * 4)
    value = flops / (batch_ms * 1e-3) / 1e9
    achieved = bytes_min / (batch_ms * 1e-3) / 1e9
    loop_ms = per_launch_s * count * 1e3
    out = {"metric": "batched small solves GFLOP/s", "value": round(value, 2), "unit": "GFLOP/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic: n log-uniform on [3,64]; general U[-1,1]+nI / SPD G G^T + I",
           "config": {"workload": "batched uneven small solves", "count": count,
                      "parallelism": "single GPU"},
           "matrices_per_s": round(count / (batch_ms * 1e-3), 1),
           "batch_ms": round(batch_ms, 4),
           "tuning": {"wg_per_cu": wg.value, "order_mode": order.value,
                      "bins_in_use": [8 << b for b in range(4) if bin_mask.value >> b & 1],
                      "waves_per_workgroup": {f"{'spd' if c % 2 else 'general'}_{8 << (c // 2)}":
                                              waves[c] for c in range(8)},
                      "tuned_ms": round(tuned.value, 4), "untuned_ms": round(untuned.value, 4)},
           "max_backward_error_sampled": worst,
           "per_matrix_launch_loop": {"ms_for_all": round(loop_ms, 2),
                                      "sample": f"{sample} one-matrix plans, one graph launch each",
                                      "speedup_of_batched": round(loop_ms / batch_ms, 1)},
           "per_matrix_entry_point_loop": {
               "ms_for_all": round(per_call_s * count * 1e3, 2),
               "sample": f"{sample2} mat_gpu_gesv_f32 / mat_gpu_posv_f32 calls",
               "speedup_of_batched": round(per_call_s * count * 1e3 / batch_ms, 1)},
           "roofline": {"kernel": "batched_kernel (one persistent kernel per size/kind class)",
                        "bound": "hbm", "achieved": round(achieved, 3), "peak": PEAK["hbm"],
                        "unit": "GB/s", "frac": round(achieved / PEAK["hbm"], 5),
                        "traffic": hbm_traffic("batched", args.profiled),
                        "avg_launch_us": round(batch_ms * 1e3, 3)}}
    return out


# ------------------------------------------------------------ drop-in, host buffers
def run_dropin(torch, libmat_amd, args):
    """What a libmat caller gets from the drop-in symbols: mat_mul,
    mat_solve and mat_solve_spd on host Mats (fp32), timed through the C ABI
    from the call to its return -- uploads, the GPU work and the downloads
    over PCIe included (the Mats are built before the timed calls, so no
    host-side copies are timed).  Median of `reps` calls after one warm-up
    call, per order n."""
    g = libmat_amd.load("f32")
    rng = np.random.default_rng(3)
    out = {}
    for n in (256, 2048, 8192):
        # each Mat points into the numpy buffer returned beside it: keep both
        A, ab = g.to_mat(rng.uniform(-1, 1, (n, n)))
        Bm, bb = g.to_mat(rng.uniform(-1, 1, (n, n)))
        Cm, cb = g.empty_mat(n, n)
        G = rng.standard_normal((n, n))
        S, sb = g.to_mat(G @ G.T / n + np.eye(n))
        b, vb = g.to_mat(rng.uniform(-1, 1, (n, 1)))
        x, xb = g.empty_mat(n, 1)
        keep = (ab, bb, cb, sb, vb, xb)
        reps = 5 if n <= 2048 else 3

        def med(fn):
            fn()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            return float(np.median(ts))

        t_mul = med(lambda: g.lib.mat_mul(C.byref(Cm), C.byref(A), C.byref(Bm)))
        t_sol = med(lambda: g.lib.mat_solve(C.byref(x), C.byref(A), C.byref(b)))
        t_spd = med(lambda: g.lib.mat_solve_spd(C.byref(x), C.byref(S), C.byref(b)))
        del keep
        out[str(n)] = {
            "mat_mul": {"ms": round(t_mul * 1e3, 3), "GFLOP/s": round(2.0 * n ** 3 / t_mul / 1e9, 2)},
            "mat_solve": {"ms": round(t_sol * 1e3, 3),
                          "GFLOP/s": round((2.0 / 3.0 * n ** 3 + 2.0 * n ** 2) / t_sol / 1e9, 2)},
            "mat_solve_spd": {"ms": round(t_spd * 1e3, 3),
                              "GFLOP/s": round((n ** 3 / 3.0 + 2.0 * n ** 2) / t_spd / 1e9, 2)}}
        log(f"drop-in n={n}: {out[str(n)]}")
    return {"metric": "drop-in end-to-end time (host Mats, PCIe transfers included)",
            "unit": "ms", "dtype": "f32", "higher_is_better": False,
            "data": "synthetic U[-1,1]; SPD G G^T / n + I",
            "config": {"workload": "libmat drop-in calls on host buffers", "orders": [256, 2048, 8192]},
            "per_order": out}


DEFAULTS = {  # n, nrhs, kappas, steps, warmup
    "mixed-lu": (16384, 16, "1e2,1e4,1e6", 2, 1), "mixed-spd": (32768, 4, "1e2,1e3", 2, 1),
    "gemm": (8192, 0, None, 20, 3), "chol": (8192, 8, None, 10, 2), "batched": (0, 1, None, 20, 3),
    "qr": (4096, 0, None, 3, 1), "dropin": (0, 0, None, 1, 0)}
# workload name of the committed HBM-traffic passes, when run at their sizes
TAGS = {"mixed-lu": "mixed_lu_n16384", "mixed-spd": "mixed_spd_n32768", "gemm": "gemm_8192",
        "chol": "chol_8192", "batched": "batched_200k", "qr": "qr_4096", "dropin": "dropin"}
# the other configurations of the default run, under these keys
ALL_CONFIGS = [("gemm", "gemm_fp32_8192"), ("chol", "chol_fp32_8192_8rhs"),
               ("mixed-spd", "mixed_spd_32768"), ("batched", "batched_200k"),
               ("dropin", "dropin_end_to_end")]


def workload_args(args, workload, steps=None, warmup=None):
    a = argparse.Namespace(**vars(args))
    dn, dr, dk, ds, dw = DEFAULTS[workload]
    a.workload = workload
    a.n = args.n or dn
    a.nrhs = args.nrhs or dr
    a.kappas = args.kappas or dk
    a.steps = steps or ds
    a.warmup = dw if warmup is None else warmup
    same = (a.n == dn and a.nrhs == dr and (workload != "batched" or a.count == 200000))
    a.profiled = TAGS[workload] if same else None
    return a


def run_one(torch, libmat_amd, a):
    if a.workload == "mixed-lu":
        return run_mixed(torch, libmat_amd, a, spd=False)
    if a.workload == "mixed-spd":
        return run_mixed(torch, libmat_amd, a, spd=True)
    if a.workload == "gemm":
        return run_gemm(torch, libmat_amd, a)
    if a.workload == "chol":
        return run_chol(torch, libmat_amd, a)
    if a.workload == "qr":
        return run_qr(torch, libmat_amd, a)
    if a.workload == "dropin":
        return run_dropin(torch, libmat_amd, a)
    return run_batched(torch, libmat_amd, a)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps of the headline (or of --workload)")
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="all",
                    choices=["all", "mixed-lu", "mixed-spd", "gemm", "chol", "batched", "qr",
                             "dropin"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--nrhs", type=int, default=None)
    ap.add_argument("--kappas", default=None)
    ap.add_argument("--passes", type=int, default=None)
    ap.add_argument("--pivoting", type=int, default=None, choices=[0, 1, 2],
                    help="mixed LU: 0 partial, 1 butterfly, 2 tournament (library default if unset)")
    ap.add_argument("--leaf", type=int, default=None,
                    help="mixed solvers: LU pivoting block width (library default if unset)")
    ap.add_argument("--count", type=int, default=200000)
    ap.add_argument("--loop-sample", type=int, default=2000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--data", default="conditioned", choices=["conditioned", "uniform"],
                    help="mixed-lu: U[-1,1] systems instead of the conditioned family (same "
                         "shapes; for counter profilers that cannot run the vendor QR)")
    ap.add_argument("--qr", default="vendor", choices=["vendor", "own"],
                    help="orthogonal factors of the conditioned systems from the vendor QR or "
                         "from this library's (for counter profilers)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-core", type=int, default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker is not None:
        cpu_worker(args.cpu_worker.split(","), args.cpu_core)
        return
    global OWN_QR
    OWN_QR = args.qr == "own"
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the "
              "multi-GPU layout is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    head = "mixed-lu" if args.workload == "all" else args.workload
    subs = ALL_CONFIGS if args.workload == "all" else []
    if args.workload == "all":
        args.n = args.nrhs = args.kappas = None
    # the CPU baselines start first, in their own pinned process, before
    # this process touches the GPU
    cpu = CpuBaselines([] if args.no_cpu else [head] + [w for w, _ in subs if w != "dropin"])

    import torch
    import libmat_amd

    torch.cuda.set_device(0)
    out = run_one(torch, libmat_amd, workload_args(args, head, args.steps, args.warmup))
    if subs:
        out["configs"] = {}
        for w, key in subs:
            log(f"config {key}")
            try:
                out["configs"][key] = run_one(torch, libmat_amd, workload_args(args, w))
            except Exception as e:  # the headline and the other configs still stand
                out["configs"][key] = {"error": repr(e)}
            torch.cuda.empty_cache()
    base = cpu.results()
    if base:
        out["cpu_baseline"] = base.get(head)
        for w, key in subs:
            if "error" not in out["configs"][key] and w in base:
                out["configs"][key]["cpu_baseline"] = base.get(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
