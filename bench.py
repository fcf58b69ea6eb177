"""Benchmark: time to a 1e-10 relative true residual for 3D Q1 variable-
coefficient diffusion at 128^3 cells (two-phase random inclusions, contrast
1e4, 30 % volume fraction, radii 3-8 cells), plain FP64 GMG-CG vs the mixed
FP32-V-cycle flexible CG, plus V-cycle GDOF/s and the HBM roofline of the
dominant kernel.

One step = one solve from x = 0 to ||b - A x|| / ||b|| < 1e-10 (FP64
verified) with the hierarchy resident in HBM; setup (GPU assembly of all
levels, transfers, colouring, coarse inverse) is timed separately.
Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 stream (Steele, Lea, Flood 2014) for reproducible inputs."""

    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))


def inclusions(n, seed=2024, frac=0.30, rmin=3.0, rmax=8.0, contrast=1e4):
    """Per-cell coefficient on n^3 cells: 1 in the matrix phase, `contrast`
    inside random balls (centres uniform in the cube, radii uniform in
    [rmin, rmax] cells) added until the inclusion volume fraction >= frac."""
    rng = SplitMix64(seed)
    inside = np.zeros((n, n, n), bool)
    c = np.arange(n) + 0.5
    target = frac * n ** 3
    count = 0
    while count < target:
        cx, cy, cz = (rng.uniform() * n for _ in range(3))
        r = rmin + (rmax - rmin) * rng.uniform()
        lo = [max(int(v - r), 0) for v in (cx, cy, cz)]
        hi = [min(int(v + r) + 1, n) for v in (cx, cy, cz)]
        X, Y, Z = np.ix_(c[lo[0]:hi[0]], c[lo[1]:hi[1]], c[lo[2]:hi[2]])
        ball = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r * r
        sub = inside[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        count += int(np.count_nonzero(ball & ~sub))
        sub |= ball
    k = np.ones((n, n, n))
    k[inside] = contrast
    return k


def spmv_bytes(nnz, N, vb):
    """algorithmic bytes of one SpMV: values + column indices, row structure,
    x read once, y written once"""
    return nnz * (vb + 4) + 4 * (N + 1) + 2 * N * vb


def sor_sweep_bytes(nnz, N, vb):
    """one <sor> sweep over all colours: matrix once, row structure, b, diag,
    x read + write, r write"""
    return nnz * (vb + 4) + 4 * (N + 1) + 5 * N * vb


def cpu_baseline(kappa, n0, levels, base, b):
    """The oracle (C restatement of femlisp's algorithm, one thread) solving
    the same system with the same plain FP64 GMG-CG."""
    from oracle import pyoracle as po
    t0 = time.perf_counter()
    h = po.Hierarchy(3, 1, n0, levels, kappa=kappa)
    t1 = time.perf_counter()
    p = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, omega=1.0, ordering=po.ORDER_COLOUR,
                  pre_steps=2, post_steps=2, base_level=base, symmetric=1, tol=1e-10,
                  maxsteps=500)
    st, x, steps, dn = h.solve(p, b)
    t2 = time.perf_counter()
    return {"setup_s": t1 - t0, "solve_s": t2 - t1, "steps": steps, "status": st}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="cells per axis (power of two)")
    ap.add_argument("--base", type=int, default=3, help="coarse-solve level (2^base cells/axis)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the fine-level SpMV / <sor> kernels and a copy of known size "
                         "(for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    from gnu_femlisp_amd import flgpu as F
    from gnu_femlisp_amd import iteration as it
    from gnu_femlisp_amd import lattice as lat
    from gnu_femlisp_amd.matlisp import GpuVector, default_context, gemm, norm
    from gnu_femlisp_amd.flgpu import check, lib

    n = args.n
    levels = int(round(np.log2(n))) + 1
    n0 = 1
    t = time.perf_counter()
    kappa = inclusions(n)
    t_gen = time.perf_counter() - t
    ctx = default_context()
    info = ctx.info()
    kv = GpuVector.from_array(kappa.ravel(), F.F64)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    log(f"[bench] device {info['arch']} cus={info['cus']}; n={n} levels={levels} "
        f"inclusion fraction {np.mean(kappa > 1):.3f} (gen {t_gen:.1f}s)")

    if args.kernels_only:
        A, b = lat.assemble(3, 1, n, (kv, n), prec=F.F64, mode=F.ASM_F64)
        A32, _ = lat.assemble(3, 1, n, (kv, n), prec=F.F32, mode=F.ASM_MFMA)
        lat.colour(A, 3, 1, n)
        lat.colour(A32, 3, 1, n)
        N = A.nrows
        for M, prec in ((A, F.F64), (A32, F.F32)):
            y = GpuVector.from_array(np.ones(N), prec)
            x = GpuVector(N, prec)
            for _ in range(args.kernel_reps):
                gemm(1.0, M, y, 0.0, x)
            bs = GpuVector.from_array(b.numpy(), prec)
            for _ in range(args.kernel_reps):
                check(lib().flgpu_sor_iterate(M.h, 1.0, 0, x.h, bs.h, y.h))
        # copy kernel of known size (64 Mi doubles = 512 MiB read + 512 MiB written)
        from gnu_femlisp_amd.matlisp import copy as vcopy
        src, dst = GpuVector(1 << 26), GpuVector(1 << 26)
        for _ in range(args.kernel_reps):
            vcopy(src, dst)
        ctx.synchronize()
        log("[bench] kernels-only done")
        return

    # ---- plain FP64: V(2,2) symmetric multicolour GS, CG -----------------------
    t = time.perf_counter()
    cs64 = it.GeometricCS(it.SOR(1.0), 2, 2, args.base, symmetric=True, prec=F.F64)
    H64 = it.Hierarchy.lattice(cs64, 3, 1, n0, levels, kappa=(kv, n), mode=F.ASM_F64)
    ctx.synchronize()
    setup_plain = time.perf_counter() - t
    A = H64.top_matrix()
    b = H64.rhs
    N, nnz = A.nrows, A.nnz
    bnorm = norm(b)
    log(f"[bench] plain setup {setup_plain:.2f}s  N={N} nnz={nnz}")

    # ---- mixed: FP32 hierarchy (BF16-MFMA element contractions) in FP64 FCG ----
    t = time.perf_counter()
    cs32 = it.GeometricCS(it.SOR(1.0), 2, 2, args.base, symmetric=True, prec=F.F32)
    H32 = it.Hierarchy.lattice(cs32, 3, 1, n0, levels, kappa=(kv, n), mode=F.ASM_MFMA)
    ctx.synchronize()
    setup_mixed = time.perf_counter() - t
    log(f"[bench] mixed setup {setup_mixed:.2f}s")

    def run(path):
        if path == "plain":
            solver = it.LinearSolver(it.CG(cs64), success_if=("<", "reduction", 1e-10),
                                     failure_if=(">", "step", 500), true_residual=True)
            bb = {"matrix": A, "rhs": b, "hierarchy": H64}
        else:
            solver = it.LinearSolver(it.FCG(cs32), success_if=("<", "reduction", 1e-10),
                                     failure_if=(">", "step", 500), true_residual=True)
            bb = {"matrix": A, "rhs": b, "hierarchy": H32}
        t0 = time.perf_counter()
        out = it.solve(solver, bb)
        dt = time.perf_counter() - t0
        return dt, out

    results = {}
    for path in ("plain", "mixed"):
        for _ in range(args.warmup):
            run(path)
        times = []
        for _ in range(args.steps):
            ctx.synchronize()
            dt, out = run(path)
            ctx.synchronize()
            times.append(dt)
            rel = out["true_residual"] / bnorm
            assert out["status"] == "success" and rel < 1e-10, (path, out["status"], rel)
        results[path] = {"times": times, "steps": out["steps"], "true_rel": rel,
                         "x": out["solution"]}
        log(f"[bench] {path}: {np.median(times):.4f}s median, {out['steps']} iterations, "
            f"true rel residual {rel:.2e}")
    x64 = results["plain"]["x"].numpy()
    xm = results["mixed"]["x"].numpy()
    mixed_vs_plain_err = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))

    # ---- V-cycle GDOF/s (one <mg-iteration> application on the top level) -------
    vcyc = {}
    for name, H, prec in (("f64", H64, F.F64), ("f32", H32, F.F64)):
        q = GpuVector(N)
        w1 = GpuVector.from_array(np.zeros(N))
        w2 = GpuVector(N)
        from gnu_femlisp_amd.matlisp import copy as vcopy
        vcopy(b, w1)
        vcopy(b, w2)
        H.apply(q, w1, w2)
        reps = 10
        ctx.timer_start()
        for _ in range(reps):
            H.apply(q, w1, w2)
        ms = ctx.timer_stop() / reps
        vcyc[name] = {"ms": ms, "gdofs": N / (ms * 1e-3) / 1e9}
        log(f"[bench] V-cycle {name}: {ms:.3f} ms, {vcyc[name]['gdofs']:.3f} GDOF/s")

    # ---- roofline: fine-level SpMV and <sor> sweep, HIP events on the library stream
    y = GpuVector.from_array(np.random.default_rng(0).uniform(-1, 1, N))
    xo = GpuVector(N)
    A32 = H32.top_matrix()
    y32 = GpuVector.from_array(y.numpy(), F.F32)
    x32 = GpuVector(N, F.F32)
    kern = {}
    for name, M, yy, xx, vb in (("spmv_f64", A, y, xo, 8), ("spmv_f32", A32, y32, x32, 4)):
        gemm(1.0, M, yy, 0.0, xx)
        ctx.timer_start()
        for _ in range(args.kernel_reps):
            gemm(1.0, M, yy, 0.0, xx)
        ms = ctx.timer_stop() / args.kernel_reps
        byt = spmv_bytes(nnz, N, vb)
        kern[name] = {"ms": ms, "bytes": byt, "gbs": byt / (ms * 1e-3) / 1e9}
    for name, M, vb, prec in (("sor_f64", A, 8, F.F64), ("sor_f32", A32, 4, F.F32)):
        xs, bs, rs = GpuVector(N, prec), GpuVector.from_array(b.numpy(), prec), GpuVector(N, prec)
        check(lib().flgpu_sor_iterate(M.h, 1.0, 0, xs.h, bs.h, rs.h))
        ctx.timer_start()
        for _ in range(args.kernel_reps):
            check(lib().flgpu_sor_iterate(M.h, 1.0, 0, xs.h, bs.h, rs.h))
        ms = ctx.timer_stop() / args.kernel_reps
        byt = sor_sweep_bytes(nnz, N, vb)
        kern[name] = {"ms": ms, "bytes": byt, "gbs": byt / (ms * 1e-3) / 1e9}
    for k, v in kern.items():
        log(f"[bench] {k}: {v['ms'] * 1e3:.1f} us, {v['gbs']:.0f} GB/s "
            f"({v['gbs'] / 8000 * 100:.1f}% of 8 TB/s)")

    # ---- CPU baseline: the oracle on the same system -----------------------------
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle, 1 thread) ...")
        cpu = cpu_baseline(kappa, n0, levels, args.base, b.numpy())
        log(f"[bench] cpu: setup {cpu['setup_s']:.1f}s solve {cpu['solve_s']:.1f}s "
            f"steps {cpu['steps']}")

    t_plain = float(np.median(results["plain"]["times"]))
    t_mixed = float(np.median(results["mixed"]["times"]))
    dom = kern["sor_f32"]
    line = {
        "metric": "s to 1e-10 rel. residual, 3D Q1 128^3; V-cycle GDOF/s; learned+mixed vs plain",
        "value": t_mixed,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64 outer / f32 V-cycle / bf16-mfma assembly",
        "data": "synthetic: two-phase random inclusions (SplitMix64 seed 2024)",
        "config": {"workload": f"3D Q1 diffusion {n}^3 cells, N={N}, nnz={nnz}, contrast 1e4, "
                               f"30% inclusions r 3-8 cells, V(2,2) multicolour GS, "
                               f"base level {args.base}",
                   "global_batch": 1, "seq_len": N, "parallelism": "single GPU"},
        "plain_fp64_s": t_plain,
        "plain_iterations": results["plain"]["steps"],
        "mixed_iterations": results["mixed"]["steps"],
        "speedup_mixed_vs_plain": t_plain / t_mixed,
        "mixed_vs_plain_solution_rel_diff": mixed_vs_plain_err,
        "true_rel_residual": {"plain": results["plain"]["true_rel"],
                              "mixed": results["mixed"]["true_rel"]},
        "setup_s": {"plain": setup_plain, "mixed": setup_mixed},
        "vcycle_gdofs": {k: v["gdofs"] for k, v in vcyc.items()},
        "vcycle_ms": {k: v["ms"] for k, v in vcyc.items()},
        "kernels_us": {k: v["ms"] * 1e3 for k, v in kern.items()},
        "roofline": {"bound": "hbm", "kernel": "k_sor_colour<float> (fine-level <sor> sweep)",
                     "achieved": dom["gbs"], "peak": 8000.0, "unit": "GB/s",
                     "frac": dom["gbs"] / 8000.0, "traffic": None},
        "roofline_spmv_f64": {"achieved": kern["spmv_f64"]["gbs"], "peak": 8000.0,
                              "frac": kern["spmv_f64"]["gbs"] / 8000.0},
        "cpu_baseline": None if cpu is None else {
            "value": cpu["solve_s"], "unit": "s", "cores": 1, "kind": "port",
            "sample": f"full plain FP64 GMG-CG solve of the same {n}^3 system "
                      f"({cpu['steps']} iterations; setup {cpu['setup_s']:.1f}s excluded)"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
