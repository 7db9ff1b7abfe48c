#!/usr/bin/env python3
"""Benchmark of the headline workload: FP64-accurate solves/s at n=16384 with the
mixed-precision LU + iterative refinement, next to the plain FP64 GPU LU, plus FP64 GEMM
fraction of the matrix-core peak.

One step = one complete FP64-accurate solve of A x = b (A resident in HBM): FP16 (or BF16)
matrix-core LU with FP32 panels, then refinement with FP64 residuals until the normwise
backward error is <= 1e-14.  A = U diag(s) V^T, U and V from QR of N(0,1) matrices,
s geometric from 1 to 1/kappa; b = A x_true with x_true uniform[-1,1].

Prints one JSON line (see README of the driver contract): value = solves/s on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6e12, "f32": 157.3e12, "f16": 2.5e15, "bf16": 2.5e15, "hbm": 8.0e12}
METRIC = "FP64-accurate solves/s, n=16384: mixed-prec LU+IR vs FP64 LU; GEMM %peak"
# the kernel the roofline line is about (the dominant kernel of the mixed solve): its launches
# are the only ones timed with HIP events inside the timed steps
ROOF_KERNEL = "getf2_panel_f32"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_system(n, kappa, nrhs, seed, torch, qr_device="cuda"):
    """A = U diag(s) V^T (QR of Gaussians), x_true uniform[-1,1], b = A x_true (FP64).
    qr_device="cpu" runs the two QRs on the host (slower; for profiler runs that cannot
    instrument the GPU QR library)."""
    dev = torch.device("cuda")
    qd = torch.device(qr_device)
    g = torch.Generator(device=qd)
    g.manual_seed(seed)
    t0 = time.time()
    U, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=qd, generator=g))
    U = U.to(dev)
    log(f"QR 1 done {time.time() - t0:.1f}s")
    V, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=qd, generator=g))
    V = V.to(dev)
    log(f"QR 2 done {time.time() - t0:.1f}s")
    g = torch.Generator(device=dev)
    g.manual_seed(seed + 1)
    s = kappa ** (-torch.arange(n, dtype=torch.float64, device=dev) / (n - 1))
    A = (U * s) @ V.T.contiguous()
    del U, V
    xt = torch.rand(n, nrhs, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    b = A @ xt
    torch.cuda.synchronize()
    log(f"system ready {time.time() - t0:.1f}s")
    return A.contiguous(), b.contiguous(), xt


def make_system_host_qr(n, kappa, nrhs, seed, ddla):
    """The same construction without torch: the two QRs by numpy (LAPACK) on the host, the
    products by the library's GEMM on the device.  Returns device matrices (A, b).  For
    profiler runs: a process where torch's own HIP runtime also runs ends, under rocprofv3, with
    a crash after the tool's finalization; processes that use only this library end cleanly."""
    rng = np.random.default_rng(seed)
    t0 = time.time()
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    log(f"QR 1 done {time.time() - t0:.1f}s")
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    log(f"QR 2 done {time.time() - t0:.1f}s")
    U *= kappa ** (-np.arange(n) / (n - 1))
    dU, dV = ddla.DeviceMatrix.from_host(U), ddla.DeviceMatrix.from_host(V)
    del U, V
    A = ddla.DeviceMatrix(n, n)
    ddla.BlasProvider.GeMM(1.0, dU, dV, 0.0, A, ddla.NO_TRANSPOSE, ddla.TRANSPOSE)
    xt = ddla.DeviceMatrix.from_host(rng.uniform(-1.0, 1.0, (n, nrhs)))
    b = ddla.DeviceMatrix(n, nrhs)
    ddla.BlasProvider.GeMM(1.0, A, xt, 0.0, b)
    ddla.synchronize()
    log(f"system ready {time.time() - t0:.1f}s")
    return A, b


def backward_error(torch, A, x, b):
    r = b - A @ x
    rn = r.abs().amax(dim=0)
    xn = x.abs().amax(dim=0)
    an = A.abs().sum(dim=1).amax()
    return float((rn / (an * xn)).max())


def time_gemm(ddla, n, reps):
    A = ddla.DeviceMatrix.random(n, n, 0xDD1A0001)
    B = ddla.DeviceMatrix.random(n, n, 0xDD1A0002)
    C = ddla.DeviceMatrix(n, n)
    ddla.BlasProvider.GeMM(1.0, A, B, 0.0, C)
    ddla.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ddla.BlasProvider.GeMM(1.0, A, B, 0.0, C)
    ddla.synchronize()
    dt = (time.perf_counter() - t0) / reps
    tf = 2.0 * n ** 3 / dt
    return {"n": n, "ms": dt * 1e3, "tflops": tf / 1e12, "frac_fp64_peak": tf / PEAK["f64"]}


def vendor_reference(torch, A_t, b_t, n):
    """The vendor libraries on the same system, through torch (reference points, not the product
    path): torch.linalg.solve (rocSOLVER getrf + getrs, FP64) on A, and the FP64 GEMM behind
    torch.matmul at 1024^3 and 8192^3, timed with HIP events."""
    def timed(f, reps):
        f()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            f()
        e.record()
        torch.cuda.synchronize()
        return s.elapsed_time(e) / reps
    out = {"library": f"torch {torch.__version__} (rocSOLVER, hipBLASLt / rocBLAS)"}
    out["linalg_solve_ms"] = timed(lambda: torch.linalg.solve(A_t, b_t), 2)
    gem = []
    for m in (1024, 8192):
        a = torch.randn(m, m, dtype=torch.float64, device="cuda")
        b = torch.randn(m, m, dtype=torch.float64, device="cuda")
        ms = timed(lambda: a @ b, 20 if m == 1024 else 3)
        tf = 2.0 * m ** 3 / ms / 1e9
        gem.append({"n": m, "ms": ms, "tflops": tf, "frac_fp64_peak": tf * 1e12 / PEAK["f64"]})
        del a, b
    out["dgemm"] = gem
    torch.cuda.empty_cache()
    return out


def measured_traffic(name, n):
    """HBM bytes per launch of a kernel group from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled against a device copy), when it covers this workload."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_n{n}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        e = json.load(f).get(name)
    return e["traffic_bytes"] if e else None


def roofline_of(name, e, n):
    """Roofline object for one profiled kernel (totals over the timed region)."""
    if e["ms"] <= 0 or e["count"] == 0:
        return None
    sec = e["ms"] / 1e3
    prec = "f16" if "f16" in name else "bf16" if "bf16" in name else "f32" if "f32" in name else "f64"
    peak_f = PEAK[prec]
    ai = e["flops"] / e["bytes"] if e["bytes"] > 0 else float("inf")
    ridge = peak_f / PEAK["hbm"]
    per_launch_ms = e["ms"] / e["count"]
    if e["flops"] > 0 and ai >= ridge:
        ach = e["flops"] / sec
        return {"kernel": name, "bound": "mfma", "achieved": ach / 1e12, "peak": peak_f / 1e12,
                "unit": "TFLOP/s", "frac": ach / peak_f, "traffic": measured_traffic(name, n),
                "avg_launch_ms": per_launch_ms, "launches": e["count"],
                "flops_per_launch": e["flops"] / e["count"]}
    ach = e["bytes"] / sec
    return {"kernel": name, "bound": "hbm", "achieved": ach / 1e9, "peak": PEAK["hbm"] / 1e9,
            "unit": "GB/s", "frac": ach / PEAK["hbm"], "traffic": measured_traffic(name, n),
            "avg_launch_ms": per_launch_ms, "launches": e["count"],
            "bytes_per_launch": e["bytes"] / e["count"]}


def batched_workload(torch, ddla, count, prec, seed):
    """The batched configuration: count problems, n uniform on 8..256, half LU on uniform[-1,1]
    entries, half SPD (G G^T + n I, G Gaussian), packed row-major in one device buffer."""
    dev = torch.device("cuda")
    rng = np.random.default_rng(seed)
    n = rng.integers(8, 257, count).astype(np.int32)
    kind = np.zeros(count, np.int32)
    kind[rng.permutation(count)[: count // 2]] = ddla.BATCH_CHOL
    off = np.zeros(count, np.int64)
    off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    total = int((n.astype(np.int64) ** 2).sum())
    dt = torch.float64 if prec == ddla.PREC_FP64 else torch.float32
    a = torch.empty(total, dtype=dt, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    for size in np.unique(n):
        for k in (ddla.BATCH_LU, ddla.BATCH_CHOL):
            idx = np.nonzero((n == size) & (kind == k))[0]
            if idx.size == 0:
                continue
            c = int(idx.size)
            if k == ddla.BATCH_LU:
                M = torch.rand(c, size, size, dtype=torch.float64, device=dev, generator=g) * 2 - 1
            else:
                G = torch.randn(c, size, size, dtype=torch.float64, device=dev, generator=g)
                M = torch.bmm(G, G.transpose(1, 2).contiguous())
                M += size * torch.eye(size, dtype=torch.float64, device=dev)
            pos = (torch.from_numpy(off[idx]).to(dev)[:, None]
                   + torch.arange(size * size, device=dev)[None, :])
            a[pos.reshape(-1)] = M.reshape(-1).to(dt)
    torch.cuda.synchronize()
    return n, kind, off, a


def bench_batched(torch, ddla, count, steps, seed, check_sample, cpu_sample, threads):
    """Problems/s of the batched factorization (one graph replay per step), FP64 and FP32,
    a sample of problems checked against the oracle, and the oracle's own rate."""
    from oracle import oracle as O
    out = {}
    for prec, name in ((ddla.PREC_FP64, "fp64"), (ddla.PREC_FP32, "fp32")):
        n, kind, off, a = batched_workload(torch, ddla, count, prec, seed)
        a0 = a.clone()
        poff = np.zeros(count, np.int64)
        poff[1:] = np.cumsum(n.astype(np.int64))[:-1]
        piv = torch.zeros(int(n.sum()), dtype=torch.int32, device="cuda")
        info = torch.zeros(count, dtype=torch.int32, device="cuda")
        plan = ddla.BatchPlan(n, kind, prec, a.data_ptr(), piv.data_ptr(), info.data_ptr(),
                              tune=True)
        desc = plan.describe()
        times = []
        for _ in range(steps):
            a.copy_(a0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan.execute()
            ddla.synchronize()
            times.append(time.perf_counter() - t0)
        ms = 1e3 * float(np.median(times))
        st = info.cpu().numpy()
        # a sample of problems against the oracle (FP64 reference on the same input)
        rng = np.random.default_rng(seed + 1)
        sample = rng.choice(count, min(check_sample, count), replace=False)
        P = piv.cpu().numpy()
        worst = 0.0
        piv_match = 0
        nlu = 0
        for i in sample:
            ni = int(n[i])
            Ai = a0[off[i]: off[i] + ni * ni].double().cpu().numpy().reshape(ni, ni).copy()
            Fi = a[off[i]: off[i] + ni * ni].double().cpu().numpy().reshape(ni, ni)
            if kind[i] == ddla.BATCH_LU:
                _, po = O.getrf(Ai)
                pi = P[poff[i]: poff[i] + ni]
                piv_match += int(np.array_equal(pi, po))
                nlu += 1
                err = np.linalg.norm(Fi - Ai) / np.linalg.norm(Ai)
            else:
                O.potrf(Ai)
                err = np.linalg.norm(np.tril(Fi) - np.tril(Ai)) / np.linalg.norm(np.tril(Ai))
            worst = max(worst, float(err))
        out[name] = {"problems": count, "ms_per_batch": ms, "problems_per_s": count / (ms / 1e3),
                     "gflops": desc["flops"] / (ms / 1e3) / 1e9,
                     "gbytes_per_s": desc["bytes"] / (ms / 1e3) / 1e9,
                     "nonzero_info": int((st != 0).sum()), "plan": desc,
                     "sample_checked": int(len(sample)),
                     "sample_max_rel_diff_vs_oracle_fp64": worst,
                     "sample_lu_pivots_identical": f"{piv_match}/{nlu}"}
        log(f"batched {name}: {ms:.2f} ms for {count} problems = {count / (ms / 1e3):.3g}/s "
            f"(nb={desc['nb']}, wg/CU={desc['wg_per_cu']}), sample max rel diff {worst:.1e}, "
            f"pivots {piv_match}/{nlu}")
        del plan, a, a0, piv, info
        torch.cuda.empty_cache()
    if cpu_sample > 0:
        # the oracle, one problem per thread, on a sample of the same distribution
        from concurrent.futures import ThreadPoolExecutor
        O.set_threads(1)
        rng = np.random.default_rng(seed + 2)
        ns = rng.integers(8, 257, cpu_sample)
        mats = []
        for j, ni in enumerate(ns):
            M = O.uniform_matrix(int(ni), int(ni), seed + 10 + j)
            mats.append((M @ M.T + ni * np.eye(ni)) if j % 2 else M)

        def run(j):
            return O.potrf(mats[j]) if j % 2 else O.getrf(mats[j])

        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(run, range(cpu_sample)))
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_sample / dt, "unit": "problems/s", "cores": threads,
                               "kind": "port",
                               "sample": f"oracle LU / Cholesky (FP64) of {cpu_sample} problems, "
                                         f"n uniform on 8..256, one problem per thread: {dt:.2f}s"}
        log("batched cpu", out["cpu_baseline"])
    return out


def cpu_threads():
    nproc = os.cpu_count() or 2
    return nproc, max(1, min(nproc // 2, 16))  # reference policy ProcessorCount/2, box share 16


def cpu_baseline(n_target, n_sample, nrhs):
    """The oracle (CPU restatement of DDLA's managed FP64 LU + solve) on a bounded sample: the
    plain FP64 LU + solve configuration itself (n = 8192, A uniform[-1,1] from seed 0xDD1A0002),
    measured, and scaled by (n_target / n_sample)^3 for the headline's unit.  Returns the
    baseline object and the sample's relative pivots (compared with the GPU's in the report)."""
    from oracle import oracle as O
    nproc, threads = cpu_threads()
    O.set_threads(threads)
    A = O.uniform_matrix(n_sample, n_sample, 0xDD1A0002)
    B = O.uniform_matrix(n_sample, nrhs, 0xDD1A0003)
    t0 = time.perf_counter()
    info, piv = O.getrf(A)
    O.getrs(A, piv, B)
    dt = time.perf_counter() - t0
    scale = (n_target / n_sample) ** 3
    return ({"value": 1.0 / (dt * scale), "unit": "solves/s", "cores": threads,
             "host_nproc": nproc, "kind": "port",
             "sample": (f"oracle FP64 blocked LU (block 128) + solve at n={n_sample} took "
                        f"{dt:.2f}s on {threads} threads (host nproc {nproc}); scaled by "
                        f"(n/{n_sample})^3={scale:.0f} to n={n_target}"),
             "measured_n": n_sample, "measured_seconds": dt}, piv)


def cpu_gemm(n, reps=3):
    """The oracle's restated managed GEMM (GeMM.cs loop nest, 6x8 AVX2 micro-kernel) at n^3."""
    from oracle import oracle as O
    nproc, threads = cpu_threads()
    O.set_threads(threads)
    A = O.uniform_matrix(n, n, 0xDD1A0001)
    B = O.uniform_matrix(n, n, 0xDD1A0001 + 1)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.matmul(A, B)
        ts.append(time.perf_counter() - t0)
    dt = float(np.median(ts))
    return {"n": n, "ms": dt * 1e3, "gflops": 2.0 * n ** 3 / dt / 1e9, "threads": threads,
            "host_nproc": nproc, "kind": "port"}


def bench_c2(ddla, n=8192, reps=2):
    """The plain FP64 LU + solve configuration on the GPU: A uniform[-1,1] (seed 0xDD1A0002,
    the oracle's stream), b = A x_true with x_true uniform[-1,1]; solves/s, backward error and
    the panel / interchange kernels' HBM rates."""
    A0 = ddla.DeviceMatrix.random(n, n, 0xDD1A0002)
    xt = ddla.DeviceMatrix.random(n, 1, 0xDD1A0003)
    b = ddla.DeviceMatrix(n, 1)
    ddla.BlasProvider.GeMM(1.0, A0, xt, 0.0, b)
    W = ddla.DeviceMatrix(n, n)
    X = ddla.DeviceMatrix(n, 1)
    times = []
    prof = {}
    for k in range(reps + 1):
        ddla.native.check(ddla.native.lib().ddla_memcpy_d2d(W.ptr, A0.ptr, n * n * 8))
        ddla.native.check(ddla.native.lib().ddla_memcpy_d2d(X.ptr, b.ptr, n * 8))
        ddla.synchronize()
        if k == reps:
            ddla.profile_reset()
            ddla.profile_enable(True)
        t0 = time.perf_counter()
        piv = ddla.PLUDecompose(W)
        ddla.PLUSolve(W, piv, X)
        ddla.synchronize()
        if k >= 1:
            times.append(time.perf_counter() - t0)
        if k == reps:
            ddla.profile_enable(False)
            prof = ddla.profile_read()
    dt = float(np.median(times[:-1] if len(times) > 1 else times))
    Ah, bh, xh = A0.to_host(), b.to_host(), X.to_host()
    be = float(np.abs(bh - Ah @ xh).max() / (np.abs(Ah).sum(axis=1).max() * np.abs(xh).max()))
    kern = {}
    for name in ("getf2_panel_f64", "laswp_list", "laswp_trsm_slab_f64", "dgemm_mfma_f64"):
        e = prof.get(name)
        if e and e["ms"] > 0:
            kern[name] = {"launches": e["count"], "ms": e["ms"],
                          "gbytes_per_s": e["bytes"] / (e["ms"] / 1e3) / 1e9,
                          "frac_hbm": e["bytes"] / (e["ms"] / 1e3) / PEAK["hbm"]}
            if e["flops"] > 0 and name.startswith("dgemm"):
                kern[name]["tflops"] = e["flops"] / (e["ms"] / 1e3) / 1e12
    return {"n": n, "ms_per_solve": dt * 1e3, "solves_per_s": 1.0 / dt,
            "gflops": (2.0 / 3.0 * n ** 3 + 2.0 * n * n) / dt / 1e9, "backward_error": be,
            "kernels": kern}, piv


def bench_ldlt_qr(ddla, n=8192, reps=2, cpu_n=None):
    """LDL^T (no pivoting) of a symmetric positive definite A = R R^T / n + I and Householder
    QR of a uniform[-1,1] n x n matrix on the GPU (median of `reps` after one warm-up), with
    the panel kernels' share; optionally the oracle's numpy restatements at cpu_n on the host."""
    lib = ddla.native.lib()
    R = ddla.DeviceMatrix.random(n, n, 0xDD1A0010)
    S = ddla.DeviceMatrix.from_host(np.eye(n))
    ddla.BlasProvider.GeMM(1.0 / n, R, R, 1.0, S, transB=ddla.TRANSPOSE)
    W = ddla.DeviceMatrix(n, n)
    T = ddla.DeviceMatrix(ddla.qr_t_rows(n, n), n)
    out = {}
    for name, src, flops, kern in (
            ("ldlt", S, n ** 3 / 3.0, "ldlt_diag_f64"),
            ("qr", R, 4.0 * n ** 3 / 3.0, "qr_panel_f64")):
        ts, prof = [], {}
        for k in range(reps + 1):
            ddla.native.check(lib.ddla_memcpy_d2d(W.ptr, src.ptr, n * n * 8))
            ddla.synchronize()
            if k == reps:
                ddla.profile_reset()
                ddla.profile_enable(True)
            t0 = time.perf_counter()
            if name == "ldlt":
                ddla.LDLTDecompose(ddla.LOWER, W)
            else:
                ddla.QRDecompose(W, T)
            ddla.synchronize()
            if k >= 1:
                ts.append(time.perf_counter() - t0)
            if k == reps:
                ddla.profile_enable(False)
                prof = ddla.profile_read()
        dt = float(np.median(ts[:-1] if len(ts) > 1 else ts))
        e = prof.get(kern, {"ms": 0.0, "count": 0})
        g = prof.get("dgemm_mfma_f64", {"ms": 0.0, "flops": 0.0})
        out[name] = {"n": n, "ms": dt * 1e3, "gflops": flops / dt / 1e9,
                     "panel_kernel": kern, "panel_ms": e["ms"], "panel_launches": e["count"],
                     "gemm_ms": g["ms"],
                     "gemm_tflops": g["flops"] / (g["ms"] / 1e3) / 1e12 if g["ms"] else None}
    if cpu_n:
        from oracle import oracle as O
        rng = np.random.default_rng(0xDD1A0011)
        M = rng.uniform(-1, 1, (cpu_n, cpu_n))
        Sc = M @ M.T / cpu_n + np.eye(cpu_n)
        t0 = time.perf_counter()
        O.ldlt(Sc)
        t_l = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.qr(M)
        t_q = time.perf_counter() - t0
        out["cpu_oracle"] = {"n": cpu_n, "ldlt_ms": t_l * 1e3, "qr_ms": t_q * 1e3,
                             "ldlt_gflops": cpu_n ** 3 / 3.0 / t_l / 1e9,
                             "qr_gflops": 4.0 * cpu_n ** 3 / 3.0 / t_q / 1e9,
                             "kind": "port",
                             "note": "numpy restatements (column loops in Python, BLAS-3 "
                                     "steps in numpy's BLAS) of LDLT.cs / QR.cs"}
    return out


def timed(fn):
    fn()
    import ddla_amd as ddla
    ddla.synchronize()
    t0 = time.perf_counter()
    r = fn()
    ddla.synchronize()
    return r, time.perf_counter() - t0


def bench_extras(torch, ddla, A, A_t, n, args, t_fp64):
    """BF16 trailing updates and 64 right-hand sides on the headline matrix; then the mixed
    Cholesky configuration: SPD A = Q diag(l) Q^T, l geometric from 1 to 1e-5, FP16 and BF16
    factors, next to the plain FP64 Cholesky on the GPU."""
    out = []
    dev = torch.device("cuda")

    def rhs(nrhs, seed):
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        xt = torch.rand(n, nrhs, dtype=torch.float64, device=dev, generator=g) * 2 - 1
        return xt, (A_t @ xt).contiguous()

    def rec(name, info, dt, xt, X_t, extra=None):
        fe = float(((X_t - xt).abs().amax() / xt.abs().amax()))
        e = {"config": name, "ms_per_solve": dt * 1e3, "iters": info.iters, "status": info.status,
             "fallback": info.fallback, "backward_error": info.backward_error,
             "forward_error_vs_xtrue": fe, "factor_ms": info.factor_ms,
             "refine_ms": info.refine_ms}
        if extra:
            e.update(extra)
        out.append(e)
        log(f"{name}: {dt * 1e3:.1f} ms iters={info.iters} berr={info.backward_error:.2e} "
            f"fallback={info.fallback}")

    for prec_name, nrhs in (("bf16", 1), ("fp16", 64)):
        xt, b_t = rhs(nrhs, args.seed + 7 + nrhs)
        X_t = torch.zeros(n, nrhs, dtype=torch.float64, device=dev)
        B = ddla.DeviceMatrix.wrap(b_t.data_ptr(), n, nrhs, nrhs, b_t)
        X = ddla.DeviceMatrix.wrap(X_t.data_ptr(), n, nrhs, nrhs, X_t)
        prec = ddla.PREC_BF16 if prec_name == "bf16" else ddla.PREC_FP16
        info, dt = timed(lambda: ddla.gesv_mpir(A, B, X, prec=prec, tol=1e-14))
        rec(f"lu {prec_name} kappa={args.kappa:g} nrhs={nrhs}", info, dt, xt, X_t,
            {"speedup_vs_fp64_lu_1rhs": t_fp64 / dt})
        del B, X, X_t, b_t, xt

    # Cholesky: one QR of a Gaussian
    g = torch.Generator(device=torch.device(args.qr_device))
    g.manual_seed(args.seed + 2)
    Q, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=args.qr_device,
                                       generator=g))
    Q = Q.to(dev)
    lam = 10.0 ** (-5.0 * torch.arange(n, dtype=torch.float64, device=dev) / (n - 1))
    S_t = ((Q * lam) @ Q.T.contiguous())
    S_t = ((S_t + S_t.T) * 0.5).contiguous()
    del Q
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed + 3)
    xt = torch.rand(n, 1, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    b_t = (S_t @ xt).contiguous()
    S = ddla.DeviceMatrix.wrap(S_t.data_ptr(), n, n, n, S_t)
    B = ddla.DeviceMatrix.wrap(b_t.data_ptr(), n, 1, 1, b_t)
    # the plain FP64 GPU Cholesky + solve on the same system
    W_t = torch.empty_like(S_t)
    Xf_t = torch.empty_like(b_t)
    W = ddla.DeviceMatrix.wrap(W_t.data_ptr(), n, n, n, W_t)
    Xf = ddla.DeviceMatrix.wrap(Xf_t.data_ptr(), n, 1, 1, Xf_t)

    def chol64():
        W_t.copy_(S_t)
        Xf_t.copy_(b_t)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ddla.CholDecompose(ddla.LOWER, W)
        ddla.CholeskySolve(ddla.LOWER, W, Xf)
        ddla.synchronize()
        return time.perf_counter() - t0

    chol64()
    t_chol64 = chol64()
    be64 = backward_error(torch, S_t, Xf_t, b_t)
    del W, W_t
    out.append({"config": "cholesky fp64 (plain GPU)", "ms_per_solve": t_chol64 * 1e3,
                "backward_error": be64})
    log(f"cholesky fp64: {t_chol64 * 1e3:.1f} ms berr={be64:.2e}")
    for prec_name in ("fp16", "bf16"):
        X_t = torch.zeros(n, 1, dtype=torch.float64, device=dev)
        X = ddla.DeviceMatrix.wrap(X_t.data_ptr(), n, 1, 1, X_t)
        prec = ddla.PREC_BF16 if prec_name == "bf16" else ddla.PREC_FP16
        info, dt = timed(lambda: ddla.posv_mpir(S, B, X, prec=prec, tol=1e-14))
        fe64 = float(((X_t - Xf_t).abs().amax() / Xf_t.abs().amax()))
        rec(f"cholesky {prec_name} lambda 1..1e-5 nrhs=1", info, dt, xt, X_t,
            {"speedup_vs_fp64_cholesky": t_chol64 / dt, "forward_error_vs_fp64_cholesky": fe64})
        del X, X_t
    del S, S_t, B, b_t, Xf, Xf_t
    torch.cuda.empty_cache()
    return out


def solve_only(ddla, args, A, B, X, n, nrhs, prec):
    """--solve-only: the mixed solver or the plain FP64 LU + solve, --warmup + --steps times."""
    piv_codes = {"partial": ddla.PIVOT_PARTIAL, "tournament": ddla.PIVOT_TOURNAMENT,
                 "butterfly": ddla.PIVOT_BUTTERFLY}
    if args.solve_only == "fp64":
        W = ddla.DeviceMatrix(n, n)
        Xf = ddla.DeviceMatrix(n, nrhs)

        def run():
            ddla.native.check(ddla.native.lib().ddla_memcpy_d2d(W.ptr, A.ptr, n * n * 8))
            ddla.native.check(ddla.native.lib().ddla_memcpy_d2d(Xf.ptr, B.ptr, n * nrhs * 8))
            piv = ddla.PLUDecompose(W)
            ddla.PLUSolve(W, piv, Xf)
            return None
    else:
        def run():
            return ddla.gesv_mpir(A, B, X, prec=prec, tol=1e-14,
                                  pivoting=piv_codes[args.pivoting])
    for _ in range(args.warmup):
        run()
    ddla.synchronize()
    t0 = time.perf_counter()
    info = None
    for _ in range(args.steps):
        info = run()
    ddla.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    out = {"solve_only": args.solve_only, "n": n, "ms_per_solve": dt * 1e3}
    if info is not None:
        out.update({"status": info.status, "iters": info.iters,
                    "backward_error": info.backward_error})
    if args.breakdown:
        # one more solve with per-launch event timing: kernel time per solve by kernel name
        ddla.profile_reset()
        ddla.profile_enable(True)
        run()
        ddla.synchronize()
        ddla.profile_enable(False)
        prof = ddla.profile_read()
        out["kernels"] = {k: {"launches": v["count"], "ms": round(v["ms"], 3)}
                          for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--prec", choices=["fp16", "bf16", "fp32"], default="fp16")
    ap.add_argument("--pivoting", choices=["partial", "tournament", "butterfly"],
                    default="partial")
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compare", default="",
                    help="comma list of other pivoting rules of the mixed solver timed once each")
    ap.add_argument("--sweep-kappa", default="1e2,1e4",
                    help="extra condition numbers timed (one solve each) besides --kappa")
    ap.add_argument("--seed", type=int, default=0xDD1A0003)
    ap.add_argument("--batched", type=int, default=20000,
                    help="problems in the batched-factorization measurement (0 skips it)")
    ap.add_argument("--qr-device", choices=["cuda", "cpu"], default="cuda",
                    help="where the test matrix's QR factors are computed")
    ap.add_argument("--gemm-only", type=int, default=0, metavar="N",
                    help="only time the FP64 GEMM C = A B at N x N x N (for profiler runs) and "
                         "print its figures")
    ap.add_argument("--solve-only", choices=["mixed", "fp64"], default=None,
                    help="only run the mixed solver or the plain FP64 LU + solve on the system "
                         "(--warmup + --steps times; for profiler runs) and print the timing")
    ap.add_argument("--breakdown", action="store_true",
                    help="with --solve-only: also time one solve per kernel (event profiling)")
    ap.add_argument("--no-vendor", action="store_true",
                    help="skip the vendor-library reference (torch.linalg.solve, torch DGEMM)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other mixed configurations (BF16, 64 right-hand sides, "
                         "Cholesky), one timed solve each")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this backend has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    # The profiler modes (--gemm-only; --solve-only with --qr-device cpu) run without torch.
    # Otherwise torch is imported before the library, the order in which both HIP runtimes
    # (torch's own and the system's) also run under rocprofv3.
    torch_free = bool(args.gemm_only) or (args.solve_only is not None and args.qr_device == "cpu")
    if not torch_free:
        import torch
    import ddla_amd as ddla

    ddla.init(0)
    if args.gemm_only:
        print(json.dumps({"gemm_fp64": time_gemm(ddla, args.gemm_only, max(args.steps, 1))}))
        return
    if torch_free:
        n, nrhs = args.n, args.nrhs
        prec = {"fp16": ddla.PREC_FP16, "bf16": ddla.PREC_BF16, "fp32": ddla.PREC_FP32}[args.prec]
        A, B = make_system_host_qr(n, args.kappa, nrhs, args.seed, ddla)
        X = ddla.DeviceMatrix(n, nrhs)
        return solve_only(ddla, args, A, B, X, n, nrhs, prec)

    n, nrhs = args.n, args.nrhs
    prec = {"fp16": ddla.PREC_FP16, "bf16": ddla.PREC_BF16, "fp32": ddla.PREC_FP32}[args.prec]
    log(f"n={n} kappa={args.kappa:g} prec={args.prec} nrhs={nrhs}")
    A_t, b_t, xt = make_system(n, args.kappa, nrhs, args.seed, torch, args.qr_device)
    A = ddla.DeviceMatrix.wrap(A_t.data_ptr(), n, n, n, A_t)
    B = ddla.DeviceMatrix.wrap(b_t.data_ptr(), n, nrhs, nrhs, b_t)
    X_t = torch.zeros(n, nrhs, dtype=torch.float64, device="cuda")
    X = ddla.DeviceMatrix.wrap(X_t.data_ptr(), n, nrhs, nrhs, X_t)
    torch.cuda.synchronize()

    piv_codes = {"partial": ddla.PIVOT_PARTIAL, "tournament": ddla.PIVOT_TOURNAMENT,
                 "butterfly": ddla.PIVOT_BUTTERFLY}
    pivoting = piv_codes[args.pivoting]

    def solve(piv):
        return ddla.gesv_mpir(A, B, X, prec=prec, tol=1e-14, pivoting=piv)

    if args.solve_only:
        return solve_only(ddla, args, A, B, X, n, nrhs, prec)

    # ---- optionally, other pivoting rules of the mixed solver (one timed solve each)
    other_out = []
    for name in [x for x in args.compare.split(",") if x and x != args.pivoting]:
        solve(piv_codes[name])
        ddla.synchronize()
        t0 = time.perf_counter()
        oinfo = solve(piv_codes[name])
        ddla.synchronize()
        t_other = time.perf_counter() - t0
        other_out.append({"pivoting": name, "ms_per_solve": t_other * 1e3, "status": oinfo.status,
                          "iters": oinfo.iters, "fallback": oinfo.fallback,
                          "backward_error": oinfo.backward_error,
                          "factor_ms": oinfo.factor_ms, "refine_ms": oinfo.refine_ms})
        log(f"mixed ({name}): {t_other * 1e3:.1f} ms status={oinfo.status} "
            f"iters={oinfo.iters} fallback={oinfo.fallback} factor={oinfo.factor_ms:.1f}ms")

    # ---- mixed precision LU + IR (the measured step)
    for w in range(args.warmup):
        info = solve(pivoting)
        log(f"warmup {w}: status={info.status} iters={info.iters} berr={info.backward_error:.2e} "
            f"factor={info.factor_ms:.1f}ms refine={info.refine_ms:.1f}ms fallback={info.fallback}")
    # the timed steps carry HIP events on the roofline kernel's launches only (its live launch
    # time); the per-kernel breakdown comes from one more solve with every launch timed
    ddla.profile_reset()
    ddla.profile_only(ROOF_KERNEL)
    ddla.profile_enable(True)
    ddla.synchronize()
    t0 = time.perf_counter()
    infos = []
    for k in range(args.steps):
        infos.append(solve(pivoting))
    ddla.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    ddla.profile_enable(False)
    prof_live = ddla.profile_read()
    ddla.profile_only(None)
    ddla.profile_reset()
    ddla.profile_enable(True)
    solve(pivoting)
    ddla.synchronize()
    ddla.profile_enable(False)
    prof = ddla.profile_read()
    last = infos[-1]
    log(f"mixed: {t_mixed * 1e3:.1f} ms/solve status={last.status} iters={last.iters} "
        f"berr={last.backward_error:.2e}")
    torch.cuda.synchronize()
    be_mixed = backward_error(torch, A_t, X_t, b_t)
    fe_mixed = float(((X_t - xt).abs().amax() / xt.abs().amax()))

    # ---- plain FP64 GPU LU + solve on the same system
    W_t = torch.empty_like(A_t)
    Xf_t = torch.empty_like(b_t)
    W = ddla.DeviceMatrix.wrap(W_t.data_ptr(), n, n, n, W_t)
    Xf = ddla.DeviceMatrix.wrap(Xf_t.data_ptr(), n, nrhs, nrhs, Xf_t)
    # timed without events (the same conditions as the mixed solve's timed steps), then one
    # more solve with every launch timed for the breakdown
    t_fp64 = []
    nfp = 1
    for k in range(args.warmup + max(1, min(args.steps, 2)) + 1):
        W_t.copy_(A_t)
        Xf_t.copy_(b_t)
        torch.cuda.synchronize()
        profiled = k == args.warmup + max(1, min(args.steps, 2))
        if profiled:
            ddla.profile_reset()
            ddla.profile_enable(True)
        t0 = time.perf_counter()
        piv = ddla.PLUDecompose(W)
        ddla.PLUSolve(W, piv, Xf)
        ddla.synchronize()
        if k >= args.warmup and not profiled:
            t_fp64.append(time.perf_counter() - t0)
    ddla.profile_enable(False)
    prof64 = ddla.profile_read()
    t_fp64 = sum(t_fp64) / len(t_fp64)
    be_fp64 = backward_error(torch, A_t, Xf_t, b_t)
    fe_vs_fp64 = float(((X_t - Xf_t).abs().amax() / Xf_t.abs().amax()))
    log(f"fp64 LU: {t_fp64 * 1e3:.1f} ms/solve berr={be_fp64:.2e}")
    del W, W_t

    # ---- FP64 GEMM (config: M=N=K=1024) and a large square GEMM
    gemm = [time_gemm(ddla, 1024, 50), time_gemm(ddla, 8192, 5)]
    log("gemm", gemm)

    # ---- batched uneven factorizations (20000 problems, FP64 and FP32)
    batched = None
    if args.batched > 0:
        nproc = os.cpu_count() or 2
        batched = bench_batched(torch, ddla, args.batched, max(3, args.steps), 0xDD1A0006,
                                check_sample=200, cpu_sample=0 if args.no_cpu else 2000,
                                threads=max(1, min(nproc // 2, 16)))

    # ---- roofline of the dominant kernel: launch times from the timed steps' own events
    dom = max(prof.items(), key=lambda kv: kv[1]["ms"])[0] if prof else None
    live = prof_live.get(dom) if dom else None
    roof = roofline_of(dom, live or prof[dom], n) if dom else None
    if roof:
        roof["launch_times_from"] = "timed steps" if live else "one profiled solve"
    breakdown = {k: {"launches": v["count"], "ms_per_step": v["ms"]}
                 for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])}
    breakdown64 = {k: {"launches": v["count"] // nfp, "ms_per_solve": v["ms"] / nfp}
                   for k, v in sorted(prof64.items(), key=lambda kv: -kv[1]["ms"])}

    # ---- the other mixed configurations at this kappa, one timed solve each: BF16 updates,
    #      64 right-hand sides, and the Cholesky + refinement configuration
    extras = None if args.no_extras else bench_extras(torch, ddla, A, A_t, n, args, t_fp64)

    # ---- the other condition numbers of the configuration, one timed solve each
    sweep = []
    for ks in [x for x in args.sweep_kappa.split(",") if x.strip()]:
        kap = float(ks)
        if kap == args.kappa:
            continue
        del A, B, X, A_t, b_t, X_t
        torch.cuda.empty_cache()
        A_t, b_t, xt = make_system(n, kap, nrhs, args.seed, torch, args.qr_device)
        A = ddla.DeviceMatrix.wrap(A_t.data_ptr(), n, n, n, A_t)
        B = ddla.DeviceMatrix.wrap(b_t.data_ptr(), n, nrhs, nrhs, b_t)
        X_t = torch.zeros(n, nrhs, dtype=torch.float64, device="cuda")
        X = ddla.DeviceMatrix.wrap(X_t.data_ptr(), n, nrhs, nrhs, X_t)
        solve(pivoting)
        ddla.synchronize()
        t0 = time.perf_counter()
        si = solve(pivoting)
        ddla.synchronize()
        ts = time.perf_counter() - t0
        sweep.append({"kappa": kap, "ms_per_solve": ts * 1e3, "solves_per_s": 1.0 / ts,
                      "iters": si.iters, "status": si.status, "backward_error": si.backward_error,
                      "speedup_vs_fp64_lu": t_fp64 / ts})
        log(f"kappa={kap:g}: {ts * 1e3:.1f} ms iters={si.iters} berr={si.backward_error:.2e}")

    # ---- the plain FP64 LU + solve configuration (n = 8192) on the GPU
    vendor = None if args.no_vendor else vendor_reference(torch, A_t, b_t, n)
    if vendor:
        vendor["speedup_of_mixed_vs_linalg_solve"] = vendor["linalg_solve_ms"] / (t_mixed * 1e3)
        vendor["speedup_of_fp64_lu_vs_linalg_solve"] = vendor["linalg_solve_ms"] / (t_fp64 * 1e3)
        log("vendor", vendor)
    c2, piv_c2 = bench_c2(ddla, 8192)
    log("fp64 LU n=8192", c2)
    ldlt_qr = None if args.no_extras else bench_ldlt_qr(ddla, 8192,
                                                       cpu_n=None if args.no_cpu else 1024)
    log("ldlt / qr", ldlt_qr)

    cpu = None
    cpu_c1 = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu, piv_cpu = cpu_baseline(n, min(args.cpu_sample, n), nrhs)
        log("cpu", cpu)
        if cpu["measured_n"] == c2["n"]:
            c2["cpu"] = {"seconds": cpu["measured_seconds"],
                         "solves_per_s": 1.0 / cpu["measured_seconds"],
                         "threads": cpu["cores"], "host_nproc": cpu["host_nproc"],
                         "kind": "port"}
            c2["pivots_identical_to_cpu_oracle"] = bool(np.array_equal(piv_c2, piv_cpu))
        cpu_c1 = cpu_gemm(1024)
        log("cpu gemm", cpu_c1)

    out = {
        "metric": METRIC,
        "value": 1.0 / t_mixed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.prec}-mfma/fp32/fp64",
        "data": "synthetic",
        "config": {"workload": f"mixed-precision LU + iterative refinement, n={n}, kappa={args.kappa:g}, "
                               f"{args.prec} trailing updates, {args.pivoting} pivoting, nrhs={nrhs}",
                   "n": n, "kappa": args.kappa, "factor_prec": args.prec, "nrhs": nrhs,
                   "pivoting": args.pivoting, "tol": 1e-14},
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed": {"iters": last.iters, "status": last.status, "fallback": last.fallback,
                  "backward_error": be_mixed, "forward_error_vs_xtrue": fe_mixed,
                  "forward_error_vs_fp64_lu": fe_vs_fp64,
                  "factor_ms": last.factor_ms, "refine_ms": last.refine_ms},
        "mixed_other_pivoting": other_out or None,
        "kappa_sweep": sweep,
        "extra_configs": extras,
        "fp64_lu": {"ms_per_solve": t_fp64 * 1e3, "solves_per_s": 1.0 / t_fp64,
                    "backward_error": be_fp64,
                    "gflops": (2.0 / 3.0 * n ** 3 + 2.0 * n * n * nrhs) / t_fp64 / 1e9},
        "speedup_vs_fp64_lu": t_fp64 / t_mixed,
        "gemm_fp64": gemm,
        "gemm_fp64_cpu_oracle": cpu_c1,
        "vendor_fp64": vendor,
        "fp64_lu_n8192": c2,
        "ldlt_qr_n8192": ldlt_qr,
        "kernel_breakdown": breakdown,
        "kernel_breakdown_fp64_lu": breakdown64,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
