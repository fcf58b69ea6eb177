"""Benchmark: time to a 1e-10 relative true residual for 3D Q1 variable-
coefficient diffusion at 128^3 cells (two-phase random inclusions, contrast
1e4, 30 % volume fraction, radii 3-8 cells):

  plain    FP64 GMG (V(2,2) multicolour GS) preconditioned CG
  mixed    FP32 V-cycle (BF16-MFMA coarse assembly) in FP64 flexible CG
  learned  mixed + the learned coarse correction (plain GMG-FCG fallback)

plus V-cycle GDOF/s and the HBM roofline of the fine-level kernels, and the
3D Q2 48^3 configuration (kappa = exp(N(0, 1.5^2)) per cell, f = 1) solved with
femlisp's default smoother for order 2, the vertex-centred block SSC.

One step = one solve from x = 0 to ||b - A x|| / ||b|| < 1e-10 (FP64
verified) with the hierarchy resident in HBM; setup (GPU assembly of all
levels, transfers, colouring, coarse inverse) and network training are timed
separately.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
WEIGHTS = os.path.join(ROOT, "gnu_femlisp_amd", "unet_weights.npz")
PATH_DESC = {
    "mixed-bb-jac": "FP32 damped-Jacobi V(2,2) V-cycle, omega 0.9, black-box transfers, Galerkin levels, "
                    "in FP64 flexible CG",
    "plain-bb-jac": "FP64 damped-Jacobi V(2,2) V-cycle, omega 0.9, black-box transfers, Galerkin levels, "
                    "in CG",
    "mixed-bb": "FP32 multicolour GS V(2,2) V-cycle, black-box transfers, in FP64 flexible CG",
    "plain-bb": "FP64 multicolour GS V(2,2) V-cycle, black-box transfers, in CG",
    "mixed": "FP32 multicolour GS V(2,2) V-cycle in FP64 flexible CG",
}
# the value path's dominant kernel as the profilers name it: the FP32 per-cell <jacobi> pass
ROOF_KERNEL = "k_q1_zblock<float, 1, unsigned int, 5, false>"


def spmv_bytes(nnz, N, vb):
    """CSR SpMV: values + column indices, row structure, x read, y written"""
    return nnz * (vb + 4) + 4 * (N + 1) + 2 * N * vb


def sor_sweep_bytes(nnz, N, vb):
    """CSR <sor> sweep: matrix once, row structure, b, diag, x read + write, r write"""
    return nnz * (vb + 4) + 4 * (N + 1) + 5 * N * vb


def cell_spmv_bytes(ncells, N, vb):
    """per-cell form SpMV: cell factors, x read, y written"""
    return ncells * vb + 2 * N * vb


def cell_sor_bytes(ncells, N, vb):
    """per-cell form <sor> sweep: cell factors, b, x read + write, r write"""
    return ncells * vb + 4 * N * vb


def cpu_baseline(kappa, n0, levels, base, b, iters=24):
    """The oracle (C restatement of femlisp's algorithm, one thread) running the
    same plain FP64 GMG-CG on the same system for a bounded sample of `iters`
    iterations (about 13 s at 128^3); the caller scales the time per iteration."""
    from oracle import pyoracle as po
    t0 = time.perf_counter()
    h = po.Hierarchy(3, 1, n0, levels, kappa=kappa)
    t1 = time.perf_counter()
    p = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, omega=1.0, ordering=po.ORDER_COLOUR,
                  pre_steps=2, post_steps=2, base_level=base, symmetric=1, tol=1e-10,
                  maxsteps=iters - 1)
    st, x, steps, dn = h.solve(p, b)
    t2 = time.perf_counter()
    return {"setup_s": t1 - t0, "sample_s": t2 - t1, "steps": steps, "status": st}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="cells per axis (power of two)")
    ap.add_argument("--base", type=int, default=3, help="coarse-solve level (2^base cells/axis)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--train-problems", type=int, default=256)
    ap.add_argument("--train-epochs", type=int, default=4)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--paths", default="plain,mixed,plain-bb,mixed-bb,plain-bb-jac,mixed-bb-jac,"
                                       "learned",
                    help="comma-separated solver paths to time")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every V-cycle kernel separately instead of replaying HIP graphs "
                         "(for kernel tracing: graph replays are not followed by every tracer)")
    ap.add_argument("--solve-only", action="store_true",
                    help="time the solves only (no V-cycle, kernel or CPU measurements; "
                         "no JSON line), e.g. under a profiler")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the fine-level kernels and a copy of known size (counter runs)")
    ap.add_argument("--no-q2", action="store_true", help="skip the Q2 48^3 configuration")
    ap.add_argument("--q2-only", action="store_true",
                    help="only the Q2 48^3 configuration (printed to stderr)")
    ap.add_argument("--q2-cycles-only", action="store_true",
                    help="only the Q2 48^3 hierarchies, V-cycles and block iterates, no solves "
                         "(for kernel tracing)")
    args = ap.parse_args()
    if not args.solve_only and "plain" not in args.paths.split(","):
        ap.error("the JSON line compares with the plain path: include it or use --solve-only")
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    from gnu_femlisp_amd import flgpu as F
    from gnu_femlisp_amd import iteration as it
    from gnu_femlisp_amd import lattice as lat
    from gnu_femlisp_amd import learned
    from gnu_femlisp_amd.fields import inclusions
    from gnu_femlisp_amd.flgpu import check, lib
    from gnu_femlisp_amd.matlisp import GpuVector, copy as vcopy, default_context, gemm, norm

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    def q2_config(reps=3, solves=True):
        """3D Q2, 48^3 cells (5 levels from 3^3), kappa = exp(g), g ~ N(0, 1.5^2) iid per cell
        (seed 48), f = 1: FP64 <geometric-ssc> V(1,1) GMG-CG (femlisp's smoother for order 2,
        geomg.lisp:58-71), the same cycle in FP32 inside FP64 FCG, and the FP32 multicolour GS
        V(2,2) cycle for comparison; then one SSC iterate on the fine level alone."""
        n0, lv, nq = 3, 5, 48
        qctx = default_context()
        kap = np.exp(np.random.default_rng(48).normal(scale=1.5, size=(nq,) * 3))
        Aq, bq = lat.assemble(3, 2, nq, kap, prec=F.F64)
        bn = norm(bq)
        out = {"workload": f"3D Q2 {nq}^3 cells, N={Aq.nrows}, nnz={Aq.nnz}, "
                           "kappa = exp(N(0, 1.5^2)) per cell (seed 48), f = 1, base 3^3"}
        hq = {}
        for name, sm, st, prec, itc in (
                ("plain-ssc", it.GeometricSSC(1.0), 1, F.F64, it.CG),
                ("mixed-ssc", it.GeometricSSC(1.0), 1, F.F32, it.FCG),
                ("mixed-sor", it.SOR(1.0), 2, F.F32, it.FCG)):
            cs = it.GeometricCS(sm, st, st, 0, symmetric=True, prec=prec)
            t0 = time.perf_counter()
            H = it.Hierarchy.lattice(cs, 3, 2, n0, lv, kappa=kap,
                                     mode=F.ASM_MFMA if prec == F.F32 else F.ASM_F64)
            qctx.synchronize()
            su = time.perf_counter() - t0
            if not args.no_graphs:
                H.use_graphs(True)
            hq[name] = H
            # V-cycle on vectors of the hierarchy's precision (a rhs from b, zero start)
            Nt = H.top_matrix().nrows
            qv, w1 = GpuVector(Nt, prec), GpuVector.from_array(bq.numpy(), prec)
            w2 = GpuVector.from_array(bq.numpy(), prec)
            H.apply(qv, w1, w2)
            qctx.timer_start()
            for _ in range(10):
                H.apply(qv, w1, w2)
            vms = qctx.timer_stop() / 10
            out[name] = {"vcycle_ms": vms, "vcycle_gdofs": Nt / (vms * 1e-3) / 1e9, "setup_s": su}
            log(f"[bench] q2 {name}: V-cycle {vms:.3f} ms")
            del qv, w1, w2
            if not solves:
                continue
            ts = []
            for _ in range(reps):
                r = it.solve(it.LinearSolver(itc(cs), success_if=("<", "reduction", 1e-10),
                                             failure_if=(">", "step", 300), true_residual=True),
                             {"matrix": Aq, "rhs": bq, "hierarchy": H})
                ts.append(r["seconds"])
            rel = r["true_residual"] / bn
            assert r["status"] == "success" and rel < 1e-10, (name, r["status"], rel)
            out[name].update({"s": float(np.median(ts)), "steps": r["steps"], "true_rel": rel})
            log(f"[bench] q2 {name}: {out[name]['s']:.4f}s, {r['steps']} iterations, "
                f"true rel residual {rel:.2e}")
        # the fine-level <geometric-ssc> iterate: 8 colour launches of k_blk_sweep.  Algorithmic
        # bytes: the CSR rows of every block's DOFs (each block recomputes its residual),
        # every block's inverse, and per DOF of a block b read, r and x written, x read
        kern = {}
        rp = Aq.csr()[0]
        keys = lat.vertex_blocks(3, 2, nq)
        present = keys >= 0
        entries = int(np.sum((rp[1:] - rp[:-1])[keys[present]]))
        for prec, vb, tag in ((F.F64, 8, "f64"), (F.F32, 4, "f32")):
            H = hq["plain-ssc" if prec == F.F64 else "mixed-ssc"]
            At = H.top_matrix()
            Nq = At.nrows
            xq = GpuVector.from_array(np.zeros(Nq), prec)
            bb = GpuVector.from_array(bq.numpy(), prec)
            rq = GpuVector(Nq, prec)
            it_ = lambda: check(lib().flgpu_block_iterate(At.h, F.SMOOTH_SSC, 1.0, 0, xq.h,
                                                          bb.h, rq.h))
            it_()
            reps_k = 10
            qctx.timer_start()
            for _ in range(reps_k):
                it_()
            us = qctx.timer_stop() * 1e3 / reps_k
            byt = entries * (4 + vb) + keys.shape[0] * 27 * 27 * vb + int(present.sum()) * 4 * vb
            kern[f"ssc_iterate_{tag}"] = {"us": us, "GBps": byt / (us * 1e-6) / 1e9,
                                          "frac": byt / (us * 1e-6) / 8e12, "bytes": byt}
            log(f"[bench] q2 ssc iterate {tag}: {us:.1f} us (8 colour launches), "
                f"{byt / (us * 1e-6) / 1e9:.0f} GB/s")
            del At
        out["kernels"] = kern
        if solves:
            out["speedup_ssc_vs_sor"] = out["mixed-sor"]["s"] / out["mixed-ssc"]["s"]
        del hq, Aq, bq
        return out

    if args.q2_only or args.q2_cycles_only:
        q2_config(solves=not args.q2_cycles_only)
        return

    n = args.n
    levels = int(round(np.log2(n))) + 1
    n0 = 1
    kappa = inclusions(n)
    ctx = default_context()
    info = ctx.info()
    kv = GpuVector.from_array(kappa.ravel(), F.F64)
    log(f"[bench] device {info['arch']} cus={info['cus']}; n={n} levels={levels} "
        f"inclusion fraction {np.mean(kappa > 1):.3f}")

    def sor(M, x, b, r):
        check(lib().flgpu_sor_iterate(M.h, 1.0, 0, x.h, b.h, r.h))

    def fine_kernels(reps):
        """(name, ms per call, algorithmic bytes) of the fine-level kernels in
        both storages and precisions"""
        out = {}
        for prec, vb, mode in ((F.F64, 8, F.ASM_F64), (F.F32, 4, F.ASM_MFMA)):
            tag = "f64" if prec == F.F64 else "f32"
            for form, fname in ((F.STORAGE_CELL, "cell"), (F.STORAGE_STENCIL, "stencil"),
                                (F.STORAGE_SELL, "sell")):
                M, b = lat.assemble(3, 1, n, (kv, n), prec=prec, mode=mode)
                assert M.storage(form) == form
                N, nnz = M.nrows, M.nnz
                y = GpuVector.from_array(np.random.default_rng(0).uniform(-1, 1, N), prec)
                x, r = GpuVector(N, prec), GpuVector(N, prec)
                # algorithmic bytes: SpMV, one <sor> sweep, one <jacobi> step (x, b read, new x
                # written; the diagonal read unless the per-cell factors give it)
                if form == F.STORAGE_CELL:
                    sb, wb = cell_spmv_bytes(n ** 3, N, vb), cell_sor_bytes(n ** 3, N, vb)
                    jb = n ** 3 * vb + 3 * N * vb
                elif form == F.STORAGE_STENCIL:
                    sb, wb = 27 * N * vb + 2 * N * vb, 27 * N * vb + 5 * N * vb
                    jb = 27 * N * vb + 4 * N * vb
                else:
                    sb, wb = spmv_bytes(nnz, N, vb), sor_sweep_bytes(nnz, N, vb)
                    jb = nnz * (vb + 4) + 4 * (N + 1) + 4 * N * vb
                # 8 <jacobi> steps per call (an even count ends in x: no copy)
                jac = (lambda: check(lib().flgpu_jacobi_steps(M.h, 0.9, 8, x.h, b.h, r.h)))
                for name, fn, byt, per in (
                        (f"{fname}_spmv_{tag}", lambda: gemm(1.0, M, y, 0.0, x), sb, 1),
                        (f"{fname}_sor_{tag}", lambda: sor(M, x, b, r), wb, 1),
                        (f"{fname}_jacobi_{tag}", jac, jb, 8)):
                    fn()
                    ctx.timer_start()
                    for _ in range(reps):
                        fn()
                    ms = ctx.timer_stop() / (reps * per)
                    out[name] = {"ms": ms, "bytes": byt, "gbs": byt / (ms * 1e-3) / 1e9}
                del M
            # the structured transfers between this level and the one below: prolongation with
            # beta 1, as the V-cycle adds the correction; restriction into its own vector.
            # Algorithmic bytes: the 27 value planes over the coarse nodes, the coarse vector,
            # the fine vector read (and, for the prolongation, written)
            P = lat.interpolation(3, 1, n // 2, prec=prec)
            Nc, Nf = P.ncols, P.nrows
            xc = GpuVector.from_array(np.random.default_rng(1).uniform(-1, 1, Nc), prec)
            yf = GpuVector.from_array(np.random.default_rng(2).uniform(-1, 1, Nf), prec)
            rc = GpuVector(Nc, prec)
            for name, fn, byt in (
                    (f"tp_prolong_{tag}", lambda: gemm(1.0, P, xc, 1.0, yf), (28 * Nc + 2 * Nf) * vb),
                    (f"tp_restrict_{tag}", lambda: gemm(1.0, P, yf, 0.0, rc, job="tn"),
                     (28 * Nc + Nf) * vb)):
                fn()
                ctx.timer_start()
                for _ in range(reps):
                    fn()
                ms = ctx.timer_stop() / reps
                out[name] = {"ms": ms, "bytes": byt, "gbs": byt / (ms * 1e-3) / 1e9}
            del P
        return out

    if args.kernels_only:
        for k, v in fine_kernels(args.kernel_reps).items():
            log(f"[bench] {k}: {v['ms'] * 1e3:.1f} us, {v['gbs']:.0f} GB/s")
        src, dst = GpuVector(1 << 26), GpuVector(1 << 26)  # 512 MiB read + 512 MiB written
        for _ in range(args.kernel_reps):
            vcopy(src, dst)
        ctx.synchronize()
        log("[bench] kernels-only done")
        return

    # ---- hierarchies ---------------------------------------------------------------
    # plain: femlisp's FP64 GMG-CG (geometric transfers, re-discretised levels);
    # mixed: FP32 V-cycle (BF16-MFMA assembled) in FP64 FCG; "-bb": the same with
    # operator-dependent (black-box) transfers and Galerkin coarse levels
    # operator-dependent (black-box) transfers and Galerkin coarse levels; "-jac": damped
    # <jacobi> V(2,2), damping 0.9, one launch per step instead of 8 colour launches per sweep
    spec = {"plain": (F.F64, F.ASM_F64, "geometric", "gs"),
            "mixed": (F.F32, F.ASM_MFMA, "geometric", "gs"),
            "plain-bb": (F.F64, F.ASM_F64, "blackbox", "gs"),
            "mixed-bb": (F.F32, F.ASM_MFMA, "blackbox", "gs"),
            "plain-bb-jac": (F.F64, F.ASM_F64, "blackbox", "jac"),
            "mixed-bb-jac": (F.F32, F.ASM_MFMA, "blackbox", "jac")}
    hier, setup = {}, {}
    paths = args.paths.split(",")
    for name, (prec, mode, tr, sm) in spec.items():
        if name not in paths and not (name == "mixed" and "learned" in paths):
            continue
        t = time.perf_counter()
        if sm == "jac":
            cs = it.GeometricCS(it.Jacobi(0.9), 2, 2, args.base, prec=prec, transfer=tr)
        else:
            cs = it.GeometricCS(it.SOR(1.0), 2, 2, args.base, symmetric=True, prec=prec,
                                transfer=tr)
        hier[name] = (cs, it.Hierarchy.lattice(cs, 3, 1, n0, levels, kappa=(kv, n), mode=mode))
        if args.no_graphs:
            hier[name][1].use_graphs(False)
        ctx.synchronize()
        setup[name] = time.perf_counter() - t
    # the outer iteration's FP64 system: an FP64 hierarchy's top level, else assembled
    f64 = [hier[k][1] for k in ("plain", "plain-bb", "plain-bb-jac") if k in hier]
    if f64:
        A, b = f64[0].top_matrix(), f64[0].rhs
    else:
        A, b = lat.assemble(3, 1, n, (kv, n), prec=F.F64, mode=F.ASM_F64)
    N, nnz = A.nrows, A.nnz
    bnorm = norm(b)
    log(f"[bench] setup " + " ".join(f"{k} {v:.2f}s" for k, v in setup.items()) +
        f"  N={N} nnz={nnz}")

    net, train_info = None, None
    if not args.no_learned and "learned" in args.paths.split(","):
        net = learned.UNet(2, 3)
        if os.path.exists(WEIGHTS):
            net.load(WEIGHTS)
            train_info = {"source": "gnu_femlisp_amd/unet_weights.npz"}
        else:
            t = time.perf_counter()
            X, T, cnt = learned.generate_pairs(n=32, problems=args.train_problems, seed=11)
            t_data = time.perf_counter() - t
            hist, vloss = learned.train(net, X, T, cnt, 32, epochs=args.train_epochs, batch=32,
                                        lr=1e-3, seed=2)
            train_info = {"pairs": cnt, "data_s": t_data,
                          "train_s": time.perf_counter() - t - t_data,
                          "loss": hist, "holdout": vloss}
            del X, T
        learned.attach(hier["mixed"][1], net, kv, n)

    def run(path):
        learned_p = path == "learned"
        cs, H = hier["mixed" if learned_p else path]
        if path.startswith("plain"):
            solver = it.LinearSolver(it.CG(cs), failure_if=(">", "step", 600), true_residual=True)
        else:
            solver = it.LinearSolver(it.FCG(cs), failure_if=(">", "step", 600),
                                     true_residual=True, learned=learned_p, stall_window=40)
        t0 = time.perf_counter()
        out = it.solve(solver, {"matrix": A, "rhs": b, "hierarchy": H})
        return time.perf_counter() - t0, out

    results = {}
    wanted = [p for p in args.paths.split(",") if p != "learned" or net is not None]
    for path in wanted:
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
        results[path] = {"s": float(np.median(times)), "steps": out["steps"], "true_rel": rel,
                         "x": out["solution"].numpy(), "fell_back": out["fell_back"]}
        log(f"[bench] {path}: {results[path]['s']:.4f}s median, {out['steps']} iterations, "
            f"true rel residual {rel:.2e}" + (" (fell back)" if out["fell_back"] else ""))
    if args.solve_only:
        log("[bench] solve-only done")
        return
    x64 = results["plain"]["x"]
    for p in results:
        results[p]["rel_diff_vs_plain"] = float(np.linalg.norm(results[p].pop("x") - x64) /
                                                np.linalg.norm(x64))

    # ---- V-cycle GDOF/s --------------------------------------------------------------
    vcyc = {}
    for name, key in (("f64", "plain"), ("f32", "mixed"), ("f64-bb", "plain-bb"),
                      ("f32-bb", "mixed-bb"), ("f64-bb-jac", "plain-bb-jac"),
                      ("f32-bb-jac", "mixed-bb-jac")):
        if key not in hier:
            continue
        H = hier[key][1]
        q, w1, w2 = GpuVector(N), GpuVector(N), GpuVector(N)
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

    # ---- roofline of the value path's dominant kernel -------------------------------
    # k_q1_zblock<float, 1, unsigned int, 5, false> (the FP32 per-cell <jacobi> pass, 5 rows per
    # thread): one damped <jacobi> step on the fine level of the mixed-bb-jac hierarchy, 4 per
    # V-cycle.  Events bracket 8 x reps back-to-back launches,
    # so the average includes the dispatch gap between launches; large copies first let the
    # host queue up ahead of the GPU.  Algorithmic bytes per launch: the cell factors, x and
    # b read once, the new x written (the diagonal follows from the factors).
    roof = None
    if "mixed-bb-jac" in hier:
        At = hier["mixed-bb-jac"][1].top_matrix()
        assert At.storage() == F.STORAGE_CELL
        xj = GpuVector(At.nrows, F.F32)
        bj = GpuVector.from_array(b.numpy(), F.F32)
        tj = GpuVector(At.nrows, F.F32)
        check(lib().flgpu_jacobi_steps(At.h, 0.9, 8, xj.h, bj.h, tj.h))
        big_a, big_b = GpuVector(1 << 26), GpuVector(1 << 26)
        for _ in range(24):
            vcopy(big_a, big_b)
        reps = 25
        ctx.timer_start()
        for _ in range(reps):
            check(lib().flgpu_jacobi_steps(At.h, 0.9, 8, xj.h, bj.h, tj.h))
        ms = ctx.timer_stop()
        del big_a, big_b
        us = ms * 1e3 / (8 * reps)
        byt = n ** 3 * 4 + 3 * At.nrows * 4  # cell factors, x and b read, new x written
        gbs = byt / (us * 1e-6) / 1e9
        # HBM bytes per dispatch of this kernel at 128^3 from rocprofv3 FETCH_SIZE (x 2.0) and
        # WRITE_SIZE (x 1.0), the factors a copy of known size gives; the command and the
        # table are in profiles/hbm_traffic_q1_128_kernels.md
        traffic = 35.7e6 if n == 128 else None
        roof = {"bound": "hbm", "kernel": f"{ROOF_KERNEL} (fine-level <jacobi> step, per-cell "
                                          f"form, {n}^3)",
                "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0,
                "traffic": traffic, "launch_us": us, "bytes_per_launch": byt}
        log(f"[bench] roofline {ROOF_KERNEL}: {us:.2f} us per launch, {gbs:.0f} GB/s "
            f"({100 * gbs / 8000:.1f}% of 8 TB/s)")
        del At, xj, bj, tj

    b_host = b.numpy()
    del hier, A, b, H
    kern = fine_kernels(args.kernel_reps)
    for k, v in kern.items():
        log(f"[bench] {k}: {v['ms'] * 1e3:.1f} us, {v['gbs']:.0f} GB/s "
            f"({v['gbs'] / 8000 * 100:.1f}% of 8 TB/s)")

    q2 = None if args.no_q2 else q2_config()

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle, 1 thread) ...")
        cpu = cpu_baseline(kappa, n0, levels, args.base, b_host)
        # time to 1e-10: the time per iteration of the sample times the iteration count the
        # same plain GMG-CG needs on the GPU to the same (true) residual
        cpu["solve_s"] = cpu["sample_s"] / max(cpu["steps"], 1) * results["plain"]["steps"]
        log(f"[bench] cpu: setup {cpu['setup_s']:.1f}s, {cpu['steps']} iterations in "
            f"{cpu['sample_s']:.1f}s -> {cpu['solve_s']:.1f}s for {results['plain']['steps']}")

    fast = min((p for p in results if p != "plain"), key=lambda p: results[p]["s"])
    dom = kern["cell_sor_f32"]
    line = {
        "metric": "s to 1e-10 rel. residual, 3D Q1 128^3; V-cycle GDOF/s; learned+mixed vs plain",
        "value": results[fast]["s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results[fast]["s"] * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64 outer / f32 V-cycle / bf16-mfma coarse assembly",
        "data": "synthetic: two-phase random inclusions (SplitMix64 seed 2024)",
        "config": {"workload": f"3D Q1 diffusion {n}^3 cells, N={N}, nnz={nnz}, contrast 1e4, "
                               f"30% inclusions r 3-8 cells, base level {args.base}; value = "
                               f"'{fast}' path ({PATH_DESC.get(fast, fast)}); plain = femlisp's "
                               f"FP64 GMG-CG, V(2,2) multicolour GS",
                   "global_batch": 1, "seq_len": N, "parallelism": "single GPU"},
        "paths": {p: {k: v for k, v in r.items()} for p, r in results.items()},
        "speedup_vs_plain": {p: results["plain"]["s"] / results[p]["s"] for p in results},
        "learned_training": train_info,
        "setup_s": setup,
        "vcycle_gdofs": {k: v["gdofs"] for k, v in vcyc.items()},
        "vcycle_ms": {k: v["ms"] for k, v in vcyc.items()},
        "kernels": {k: {"us": v["ms"] * 1e3, "GBps": v["gbs"], "frac": v["gbs"] / 8000.0}
                    for k, v in kern.items()},
        "q2_48": q2,
        "roofline": roof if roof is not None else {
            "bound": "hbm", "kernel": "k_q1_sor<float> (fine-level <sor> sweep, per-cell form)",
            "achieved": dom["gbs"], "peak": 8000.0, "unit": "GB/s",
            "frac": dom["gbs"] / 8000.0, "traffic": None},
        "cpu_baseline": None if cpu is None else {
            "value": cpu["solve_s"], "unit": "s", "cores": 1, "kind": "port",
            "sample": f"plain FP64 GMG-CG on the same {n}^3 system, {cpu['steps']} iterations "
                      f"in {cpu['sample_s']:.1f}s, scaled to the {results['plain']['steps']} "
                      f"iterations of the plain path (setup {cpu['setup_s']:.1f}s excluded)"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
