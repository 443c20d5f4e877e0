"""Benchmark of the FP64 LU-solve path on one MI355X.

Workload: plain FP64 blocked LU with partial pivoting and solve, A ~ N(0,1)
n x n with no diagonal shift, n = 16384, one right-hand side ~ N(0,1).  One
step = OxiBLAS.lu(A) (copy + getrf on the GPU) followed by F \\ b, inputs
resident in HBM.  value = (2/3 n^3 + 2 n^2) / seconds per step, in GFLOP/s.

Also reported (extra keys): the trailing-update GEMM roofline measured with
HIP events on the stream it runs on, the standalone GEMM shapes, the 64-RHS
solve, the mixed-precision LU + refinement against the FP64 LU-solve
(A = U diag(sigma) V^T with Haar U, V, generated on the GPU and not timed;
speedup = FP64 LU-solve time / mixed solve time, the mixed solve run to a
componentwise backward error <= 1e-15 n), the oracle's (n = 8192) and
scipy-OpenBLAS's (n = 16384) LU-solve on the host cores, rocSOLVER's getrf
and the Cholesky factorization at the same size, and a parity check of the
timed factorization.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 16384] [--nb 0]
                  [--no-cpu] [--no-extra] [--no-mixed] [--kappa 1e6]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oxiblas.jl_amd"))

PEAK_FP64_TFLOPS = 78.6     # MI355X published dense FP64 matrix peak
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gemm_standalone(ox, torch, m, n, k, reps=5):
    """TFLOP/s of oxi_dgemm at one shape (alpha=1, beta=-1), HIP events on the
    handle's stream (torch's current stream)."""
    A = ox.DMatrix(torch.randn(m * k, dtype=torch.float64, device="cuda"), m, k)
    B = ox.DMatrix(torch.randn(k * n, dtype=torch.float64, device="cuda"), k, n)
    C = ox.DMatrix(torch.randn(m * n, dtype=torch.float64, device="cuda"), m, n)
    ox.gemm_(C, A, B, 1.0, -1.0)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ox.gemm_(C, A, B, 1.0, -1.0)
    e1.record(); torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del A, B, C
    return 2.0 * m * n * k / (ms * 1e-3) / 1e12, ms


def lowp_update_times(ox, torch, n, reps=10):
    """The mixed LU's low-precision trailing product at its update shape:
    C (n x n, FP32) -= round(A) (n x 256) round(B), per low type, operand
    rounding included.  Returns {type: (ms, TFLOP/s, fraction of the dense
    peak, C bandwidth in GB/s)}; C is read and written once (8 n^2 bytes)."""
    g = torch.Generator(device="cuda").manual_seed(0x10F)
    A = ox.DMatrix(torch.randn(n * 256, dtype=torch.float32, device="cuda", generator=g), n, 256)
    B = ox.DMatrix(torch.randn(256 * n, dtype=torch.float32, device="cuda", generator=g), 256, n)
    C = ox.DMatrix(torch.randn(n * n, dtype=torch.float32, device="cuda", generator=g), n, n)
    peak = {"bfloat16": 2.5e15, "float16": 2.5e15, "float8_e4m3fn": 5.0e15}
    out = {}
    for low in ("bfloat16", "float16", "float8_e4m3fn"):
        for _ in range(5):          # warm: the first calls also bring up the clocks
            ox.gemm_lowp_update_(C, A, B, lowtype=low)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ox.gemm_lowp_update_(C, A, B, lowtype=low)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        tf = 2.0 * n * n * 256 / t
        out[low] = (t * 1e3, tf / 1e12, tf / peak[low], 8.0 * n * n / t / 1e9)
    del A, B, C
    return out


def haar_pair(torch, n, seed):
    """U, V: the Q factors of Gaussian matrices with the sign of diag(R) fixed
    (Haar-distributed orthogonal matrices; test data only)."""
    g = torch.Generator(device="cuda").manual_seed(seed)

    def q():
        Q, R = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
        return Q * torch.sign(torch.diagonal(R))

    return q(), q()


def mixed_extra(ox, torch, n, kappa, t_fp64, seed=0x0B1A5):
    """Mixed LU + refinement on A = U diag(sigma) V^T, sigma_i =
    kappa^(-(i-1)/(n-1)), per low type: time, speedup over the FP64 LU-solve,
    iterations, errors.  The refinement stops at the target's FP64 backward
    error: componentwise backward error <= 1e-15 n.  At the headline kappa
    every low type runs; at kappa = 1e2 the low-precision trailing products,
    the only condition at which their factors still precondition GMRES-IR at
    this size.  Runs that cannot converge are bounded (maxiter 4)."""
    U, V = haar_pair(torch, n, seed)
    g = torch.Generator(device="cuda").manual_seed(seed + 1)
    xt = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    rtol = 1e-15 * n
    out = {}
    # (low type, solver, pivot); "tournament": the FP32 factor's leaves pick
    # their pivots by a tournament (one exchange per 32 columns)
    plan = ((kappa, (("float32", "ir", "partial"), ("float32", "gmres", "partial"),
                     ("float32", "gmres", "tournament"), ("bfloat16", "gmres", "partial"),
                     ("float16", "gmres", "partial"), ("float8_e4m3fn", "gmres", "partial"))),
            (1e2, (("bfloat16", "gmres", "partial"), ("float16", "gmres", "partial"),
                   ("float8_e4m3fn", "gmres", "partial"))))
    for kap, runs in plan:
        s = kap ** (-torch.arange(n, dtype=torch.float64, device="cuda") / max(n - 1, 1))
        At = (U * s) @ V.T
        b = At @ xt
        A = ox.DMatrix(At.T.contiguous().reshape(-1), n, n)   # column-major A
        bd = ox.DMatrix(b.clone(), n, 1)
        del At
        for low, solver, pivot in runs:
            # runs that measurements show cannot converge here are bounded
            expected = low == "float32" or (low == "float16" and kap <= 1e2)
            kw = dict(lowtype=low, solver=solver, maxiter=20 if expected else 4, rtol=rtol, gmres_restart=30,
                      pivot=pivot)
            r = ox.mixed_solve(A, bd, **kw)   # warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = ox.mixed_solve(A, bd, **kw)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            x = r.x.buf[:n]
            fwd = (torch.linalg.norm(x - xt) / torch.linalg.norm(xt)).item()
            key = f"mixed_{low}_{solver}{'_tournament' if pivot != 'partial' else ''}_kappa{kap:g}"
            out[key] = {"kappa": kap, "ms": round(t * 1e3, 2), "speedup_vs_fp64_lu_solve": round(t_fp64 / t, 3),
                        "converged": r.converged, "backward_error": r.backward_error,
                        "backward_error_over_n": r.backward_error / n, "forward_error": fwd,
                        "iterations": r.iterations, "gmres_iterations": r.inner_iterations, "fallback": r.fallback,
                        "factor_ms": round(r.factor_ms, 2), "refine_ms": round(r.refine_ms, 2),
                        "maxiter": kw["maxiter"], "pivot": pivot}
            log(f"mixed {low}/{solver}/{pivot} n={n} kappa={kap:g}: {t * 1e3:.1f} ms ({t_fp64 / t:.2f}x), "
                f"converged={r.converged} berr {r.backward_error:.2e}, fwd {fwd:.2e}, "
                f"it {r.iterations}/{r.inner_iterations}, factor {r.factor_ms:.1f} ms refine {r.refine_ms:.1f} ms")
        del A, bd, b
    del U, V
    return out


def cpu_baseline(n=8192, reps=3, n_openblas=16384):
    """The oracle's restatement of lu_factor + solve on all host cores, and
    scipy-OpenBLAS beside it (both labelled proxies for the Julia reference,
    which cannot run in this environment).  The oracle runs a bounded sample
    (n = 8192, about 13 s a run on 16 cores: the headline n = 16384 takes
    about 100 s a run), after a warm-up at n = 2048; OpenBLAS runs at the
    headline size."""
    import numpy as np
    import oracle as O
    import scipy.linalg as sl
    cores = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(0x0B1A5)
    fl = lambda m: 2.0 / 3.0 * m ** 3 + 2.0 * m ** 2   # noqa: E731
    O.lib()
    W = rng.standard_normal((2048, 2048))
    F, ipiv, info = O.dgetrf(W, 128, threads=cores)          # warm-up (threads, pages)
    A = rng.standard_normal((n, n)); b = rng.standard_normal(n)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        F, ipiv, info = O.dgetrf(A, 128, threads=cores)
        O.dgetrs(F, ipiv, b, threads=cores)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    del F, A
    os.environ.setdefault("OPENBLAS_NUM_THREADS", str(cores))
    A = rng.standard_normal((n_openblas, n_openblas)); b = rng.standard_normal(n_openblas)
    sl.lu_factor(W, check_finite=False)                       # warm-up
    ob = []
    for _ in range(reps):
        t0 = time.perf_counter()
        lu, piv = sl.lu_factor(A, check_finite=False)
        sl.lu_solve((lu, piv), b, check_finite=False)
        ob.append(time.perf_counter() - t0)
    to = sorted(ob)[len(ob) // 2]
    del A, lu
    # the 512^3 GEMM with beta = 0 (the reference's own CPU-sized case): the
    # oracle's packed gemm! on all host cores, median of 5 warm runs
    Ag = rng.standard_normal((512, 512)); Bg = rng.standard_normal((512, 512))
    Cg = np.asfortranarray(np.zeros((512, 512)))
    O.dgemm(Cg, Ag, Bg, 1.0, 0.0, threads=cores)
    gt = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.dgemm(Cg, Ag, Bg, 1.0, 0.0, threads=cores)
        gt.append(time.perf_counter() - t0)
    tg = sorted(gt)[2]
    return ({"value": round(fl(n) / t / 1e9, 2), "unit": "GFLOP/s", "cores": cores, "kind": "port",
             "sample": f"oracle lu_factor(blocksize=128)+solve, n={n}, 1 rhs, median of {reps} warm runs, "
                       f"{cores} threads (n=16384 takes ~100 s a run)"},
            {"value": round(fl(n_openblas) / to / 1e9, 2), "unit": "GFLOP/s", "threads": cores,
             "sample": f"scipy-OpenBLAS dgetrf+dgetrs n={n_openblas}, median of {reps} warm runs",
             "oracle_gemm_512_gflops": round(2.0 * 512 ** 3 / tg / 1e9, 2),
             "oracle_gemm_512_sample": f"oracle gemm! 512^3, beta 0, {cores} threads, median of 5 warm runs"})


def trailing_shapes(n=16384):
    """(M, N, K) in the factorization's transposed coordinates of the trailing
    updates that run on the 128 x 128 kernel, in issue order: getrf's tuned
    block-size schedule (512 while more than 3/8 of the columns remain, then
    256), its three updates per panel [j0, j1) (the next panel's columns
    [j1, j2) on the side stream, then [j2, j3) and [j3, n) on the main
    stream) and the GEMM's tile rule (128 x 128 tiles when that grid has at
    least 448 workgroups)."""
    cdiv = lambda a, b: (a + b - 1) // b
    wide_until = n - (3 * n) // 8
    nb_at = lambda j: 512 if j < wide_until else 256
    out, j0 = [], 0
    while j0 < n:
        j1 = min(j0 + nb_at(j0), n)
        if j1 >= n:
            break
        j2 = min(j1 + nb_at(j1), n)
        j3 = min(j2 + nb_at(j2), n) if j2 < n else j2
        for (M, N) in ((j2 - j1, n - j1), (j3 - j2, n - j1), (n - j3, n - j1)):
            if M > 0 and cdiv(M, 128) * cdiv(N, 128) >= 448:
                out.append((M, N, j1 - j0))
        j0 = j1
    return out


def traffic_summary(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, out_json):
    """HBM bytes per dgemm_m16_kernel launch of the n = 16384 factorization
    from a FETCH_SIZE run and a WRITE_SIZE run of
    'bench.py --steps 1 --warmup 0 --no-cpu --no-extra --no-mixed', each
    scaled by what the same counter reads for a 1 GiB device copy
    ('bench.py --copy-probe 1024')."""
    import csv

    def rows(path, counter, name):
        return [float(r["Counter_Value"]) for r in csv.DictReader(open(path))
                if r["Counter_Name"] == counter and name in r["Kernel_Name"]]
    f = rows(fetch_csv, "FETCH_SIZE", "dgemm_m16_kernel<false, false, false, 0>")
    w = rows(write_csv, "WRITE_SIZE", "dgemm_m16_kernel<false, false, false, 0>")
    cf = rows(copy_fetch_csv, "FETCH_SIZE", "copyBuffer")
    cw = rows(copy_write_csv, "WRITE_SIZE", "copyBuffer")
    kf = (1 << 20) / (sum(cf) / len(cf))          # KiB per KiB the counter reports
    kw = (1 << 20) / (sum(cw) / len(cw))
    shapes = trailing_shapes()
    reps = len(f) // len(shapes)
    assert reps >= 1 and len(f) == reps * len(shapes) and len(w) == len(f), (len(f), len(w), len(shapes))
    alg = [8 * (M * K + K * N + 2 * M * N) for (M, N, K) in shapes] * reps
    meas = [(a * kf + b * kw) * 1024 for a, b in zip(f, w)]
    out = {"kernel": "dgemm_m16_kernel: getrf trailing updates on 128x128 tiles, n=16384, tuned block sizes",
           "hbm_bytes_per_launch": round(sum(meas) / len(meas)),
           "algorithmic_bytes_per_launch": round(sum(alg) / len(alg)),
           "fetch_scale": round(kf, 4), "write_scale": round(kw, 4),
           "launches_measured": len(meas), "launches_per_factorization": len(shapes)}
    with open(out_json, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


def run_part(ox, torch, part, n, reps=3):
    """One extra workload, reps times, synchronised, with its time per run on
    stderr: the command the kernel summaries under profiles/ come from."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(0x0B1A5)
    if part == "gemm":
        # the LU's trailing-update shape on the 128 x 128 kernel: C (16384 x
        # 16384) -= A (16384 x 256) B (256 x 16384), as getrf issues it
        A = ox.DMatrix(torch.randn(n * 256, dtype=torch.float64, device=dev, generator=g), n, 256)
        B = ox.DMatrix(torch.randn(256 * n, dtype=torch.float64, device=dev, generator=g), 256, n)
        C = ox.DMatrix(torch.randn(n * n, dtype=torch.float64, device=dev, generator=g), n, n)
        work = lambda: ox.gemm_(C, A, B, -1.0, 1.0)             # noqa: E731
        reps = 10
    elif part == "lowp_gemm":
        for low, (ms, tf, frac, gbs) in lowp_update_times(ox, torch, n).items():
            log(f"lowp_gemm {low} {n}x{n}x256: {ms:.3f} ms {tf:.1f} TFLOP/s "
                f"({100 * frac:.1f} % of the dense peak), C at {gbs:.0f} GB/s")
        return
    elif part == "cholesky":
        S = ox.DMatrix(torch.randn(n * n, dtype=torch.float64, device=dev, generator=g), n, n)
        St = S.buf.view(n, n)
        St.copy_(St @ St.T)
        St.diagonal().add_(n)
        work = lambda: ox.cholesky_(S.copy())                     # noqa: E731
    elif part in ("getrs64", "trsv"):
        A = ox.DMatrix(torch.randn(n * n, dtype=torch.float64, device=dev, generator=g), n, n)
        F = ox.lu(A)
        k = 64 if part == "getrs64" else 1
        B = ox.DMatrix(torch.randn(n * k, dtype=torch.float64, device=dev, generator=g), n, k)
        work = lambda: F.ldiv_(B.copy())                          # noqa: E731
    elif part in ("mixed32", "mixed32tp", "lowp"):
        U, V = haar_pair(torch, n, 0x0B1A5)
        kap = 1e2 if part == "lowp" else 1e6
        s = kap ** (-torch.arange(n, dtype=torch.float64, device=dev) / max(n - 1, 1))
        At = (U * s) @ V.T
        del U, V
        b = At @ torch.randn(n, dtype=torch.float64, device=dev, generator=g)
        A = ox.DMatrix(At.T.contiguous().reshape(-1), n, n)
        del At
        bd = ox.DMatrix(b, n, 1)
        low = "bfloat16" if part == "lowp" else "float32"
        pivot = "tournament" if part == "mixed32tp" else "partial"
        work = lambda: ox.mixed_solve(A, bd, lowtype=low, solver="gmres", rtol=1e-15 * n, maxiter=20,  # noqa: E731
                                      pivot=pivot)
    work()
    torch.cuda.synchronize()
    for _ in range(reps):
        t0 = time.perf_counter()
        work()
        torch.cuda.synchronize()
        log(f"{part} n={n}: {(time.perf_counter() - t0) * 1e3:.3f} ms")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nb", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--traffic-summary", nargs=5, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV",
                                                           "COPY_WRITE_CSV", "OUT_JSON"),
                    help="write the HBM traffic per dgemm_m16_kernel launch from counter runs and exit")
    ap.add_argument("--part", choices=("cholesky", "getrs64", "trsv", "mixed32", "mixed32tp", "lowp", "gemm", "lowp_gemm"),
                    default=None,
                    help="only run one extra part a few times (for rocprofv3 summaries of its kernels) and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (5 times): the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.traffic_summary:
        traffic_summary(*args.traffic_summary)
        return
    if args.copy_probe:
        nel = args.copy_probe * (1 << 20) // 8
        src = torch.randn(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": nel * 8, "copy_probe_bytes_written": nel * 8, "launches": 5}))
        return
    import oxiblas_mi355x as ox
    ox.load()
    n = args.n
    if args.part:
        run_part(ox, torch, args.part, n)
        return
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(0x0B1A5)
    A0 = ox.DMatrix(torch.randn(n * n, dtype=torch.float64, device=dev, generator=g), n, n)
    b0 = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    flops = 2.0 / 3.0 * n ** 3 + 2.0 * n ** 2

    def step():
        F = ox.lu(A0, check=True, blocksize=args.nb)          # copy + getrf
        x = ox.DMatrix(b0.clone(), n, 1)
        F.ldiv_(x)                                             # F \ b
        return F, x

    for _ in range(args.warmup):
        F, x = step()
        del F, x
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        F, x = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"lu+solve n={n}: {dt * 1e3:.2f} ms/step, {flops / dt / 1e9:.1f} GFLOP/s")

    # parity of the last timed factorization: ||Ax - b|| / (||A|| ||x|| n eps)
    At = A0.buf.view(n, n)
    xt = x.buf[:n]
    r = At.T @ xt - b0
    eps = 2.0 ** -52
    berr = torch.linalg.norm(r).item() / (torch.linalg.matrix_norm(At).item() * torch.linalg.norm(xt).item() * n * eps)
    del F, x

    # roofline of the dominant kernel: the trailing-update GEMM inside getrf on
    # its 128 x 128 tiles (dgemm_m16_kernel), events around each of those
    # launches on the stream it runs on
    h = ox.handle()
    h._L.oxi_profile_enable(h.ptr, 1)
    F, x = step()
    import ctypes as C
    ms = C.c_double(); fl = C.c_double(); nl = C.c_int64()
    h._L.oxi_profile_read(h.ptr, C.byref(ms), C.byref(fl), C.byref(nl))
    h._L.oxi_profile_enable(h.ptr, 0)
    del F, x
    gemm_tflops = fl.value / (ms.value * 1e-3) / 1e12 if ms.value > 0 else 0.0
    log(f"trailing GEMM (dgemm_m16_kernel): {nl.value} launches, {ms.value:.2f} ms, {gemm_tflops:.1f} TFLOP/s")

    extra = {}
    if not args.no_extra:
        for (m_, n_, k_) in ((8192, 8192, 8192), (16384, 16384, 256)):
            tf, gms = gemm_standalone(ox, torch, m_, n_, k_)
            extra[f"gemm_{m_}x{n_}x{k_}_tflops"] = round(tf, 2)
            log(f"gemm {m_}x{n_}x{k_}: {gms:.3f} ms {tf:.1f} TFLOP/s")
        for low, (ms, tf, frac, gbs) in lowp_update_times(ox, torch, n).items():
            extra[f"lowp_update_{low}_{n}x{n}x256"] = {"ms": round(ms, 3), "tflops": round(tf, 1),
                                                       "frac_dense_peak": round(frac, 4), "c_gbs": round(gbs)}
            log(f"lowp update {low} {n}x{n}x256: {ms:.3f} ms {tf:.1f} TFLOP/s, C at {gbs:.0f} GB/s")
        F, _ = step()
        B64 = ox.DMatrix(torch.randn(n * 64, dtype=torch.float64, device=dev, generator=g), n, 64)
        B64b = B64.copy()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        F.ldiv_(B64)                 # the first solve makes the factors' solve plan
        torch.cuda.synchronize()
        extra["getrs_64rhs_first_ms"] = round((time.perf_counter() - t1) * 1e3, 3)
        t1 = time.perf_counter()
        F.ldiv_(B64b)
        torch.cuda.synchronize()
        ts = time.perf_counter() - t1
        extra["getrs_64rhs_ms"] = round(ts * 1e3, 3)
        extra["getrs_64rhs_gflops"] = round(2.0 * n * n * 64 / ts / 1e9, 1)
        log(f"getrs 64 rhs n={n}: {ts * 1e3:.3f} ms (first solve, with the plan: {extra['getrs_64rhs_first_ms']} ms)")
        # a repeated one-RHS solve on the same factors reuses their solve plan
        # (made by the factors' second solve, its one-RHS part by the first
        # one-RHS solve through it)
        b1 = ox.DMatrix(b0.clone(), n, 1)
        F.ldiv_(b1.copy())
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        F.ldiv_(b1)
        torch.cuda.synchronize()
        extra["getrs_1rhs_repeat_ms"] = round((time.perf_counter() - t1) * 1e3, 3)
        del F, B64, B64b, b1
        # the pullback of X = A \ B at n = 8192 with 64 right-hand sides
        # (A = N(0,1) + n I): Bbar by the transposed solve on the factors,
        # Abar = -Bbar X^T
        na = 8192
        Aa = ox.DMatrix(torch.randn(na * na, dtype=torch.float64, device=dev, generator=g), na, na)
        Aa.buf.view(na, na).diagonal().add_(float(na))
        Fa = ox.lu(Aa)
        Ba = ox.DMatrix(torch.randn(na * 64, dtype=torch.float64, device=dev, generator=g), na, 64)
        Xa = Fa.solve(Ba)
        Xbar = ox.DMatrix(torch.randn(na * 64, dtype=torch.float64, device=dev, generator=g), na, 64)
        ox.ldiv_pullback(Fa, Xa, Xbar)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ox.ldiv_pullback(Fa, Xa, Xbar)
        torch.cuda.synchronize()
        extra["adjoint_ldiv_n8192_64rhs_ms"] = round((time.perf_counter() - t1) * 1e3, 3)
        log(f"adjoint of A\\B n={na}, 64 rhs: {extra['adjoint_ldiv_n8192_64rhs_ms']} ms")
        del Aa, Fa, Ba, Xa, Xbar
        # rocSOLVER's getrf through torch on the same box, for comparison (labelled)
        Ac = A0.buf.view(n, n).clone()
        torch.linalg.lu_factor(Ac)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(3):
            torch.linalg.lu_factor(Ac)
        torch.cuda.synchronize()
        tr = (time.perf_counter() - t1) / 3
        extra["rocsolver_getrf_ms"] = round(tr * 1e3, 2)
        log(f"rocSOLVER getrf (torch.linalg.lu_factor) n={n}: {tr * 1e3:.2f} ms")
        del Ac
        # Cholesky of an SPD matrix of the same size (potrf2 recursion)
        S = ox.DMatrix(torch.randn(n * n, dtype=torch.float64, device=dev, generator=g), n, n)
        St = S.buf.view(n, n)
        St.copy_(St @ St.T)
        St.diagonal().add_(n)
        C0 = S.copy()
        ox.cholesky_(C0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        C1 = S.copy()
        ox.cholesky_(C1)
        torch.cuda.synchronize()
        tc = time.perf_counter() - t1
        extra["cholesky_ms"] = round(tc * 1e3, 2)
        extra["cholesky_tflops"] = round(n ** 3 / 3.0 / tc / 1e12, 2)
        log(f"cholesky n={n}: {tc * 1e3:.2f} ms")
        del S, St, C0, C1

    mixed = None
    if not args.no_mixed:
        mixed = mixed_extra(ox, torch, n, args.kappa, dt)

    cpu = None
    cpu_ob = None
    if not args.no_cpu:
        cpu, cpu_ob = cpu_baseline()
        log(f"cpu baseline: {cpu}")

    value = flops / dt / 1e9
    out = {
        "metric": "FP64 LU-solve GFLOP/s & %MFMA peak n=16384; mixed LU+IR speedup at FP64 accuracy",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: A ~ N(0,1) n x n (no diagonal shift), b ~ N(0,1), fixed seed, generated on the GPU",
        "config": {"workload": f"FP64 LU with partial pivoting + solve (1 rhs), n={n}", "n": n,
                   "blocksize": args.nb or "tuned", "parallelism": "single GPU"},
        "pct_fp64_peak": round(100.0 * value / (PEAK_FP64_TFLOPS * 1e3), 2),
        "roofline": {"bound": "mfma", "achieved": round(gemm_tflops, 3), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(gemm_tflops / PEAK_FP64_TFLOPS, 4), "traffic": None,
                     "kernel": "dgemm_m16_kernel: getrf trailing updates on 128x128 tiles (all such launches of one step)",
                     "launches": int(nl.value)},
        "cpu_baseline": cpu,
        "cpu_openblas": cpu_ob,
        "parity": {"backward_error_over_n_eps": berr, "bound": 30, "ok": bool(berr <= 30)},
    }
    # HBM bytes per trailing-update launch, from the FETCH_SIZE / WRITE_SIZE counter runs of this
    # workload (scaled by a plain-copy calibration); measured only at n=16384 with the tuned blocksize
    tf = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "lu_n16384_trailing_dgemm_traffic.json")
    if n == 16384 and not args.nb and os.path.exists(tf):
        t = json.load(open(tf))
        out["roofline"]["traffic"] = t["hbm_bytes_per_launch"]
        out["roofline"]["traffic_algorithmic"] = t["algorithmic_bytes_per_launch"]
        out["roofline"]["traffic_launches"] = f"{t['launches_measured']} dgemm_m16_kernel launches"
    if mixed:
        at = [v for v in mixed.values() if v["kappa"] == args.kappa and v["converged"]]
        best = max(at, key=lambda v: v["speedup_vs_fp64_lu_solve"]) if at else None
        out["mixed"] = {"kappa": args.kappa, "n": n, "target_backward_error": 1e-15 * n,
                        "best_speedup_vs_fp64_lu_solve": best["speedup_vs_fp64_lu_solve"] if best else None,
                        "runs": mixed}
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
