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
    "learned": "mixed-bb-jac with the learned correction: the U-Net, trained on 32^3 random-"
               "coefficient problems, maps the residual left by the V-cycle, restricted to the "
               "32^3 level, to a correction there, prolongated and added along the energy-optimal "
               "step inside FP64 flexible CG; plain GMG-FCG takes over on stall",
    "mixed-bb-jac": "FP32 damped-Jacobi V(2,2) V-cycle, omega 0.9, black-box transfers, Galerkin levels "
                    "down to 16^3 and an exact solve there, in FP64 flexible CG",
    "plain-bb-jac": "FP64 damped-Jacobi V(2,2) V-cycle, omega 0.9, black-box transfers, Galerkin levels "
                    "down to 16^3 and an exact solve there, in CG",
    "mixed-bb": "FP32 multicolour GS V(2,2) V-cycle, black-box transfers, exact 16^3 solve, in FP64 "
                "flexible CG",
    "plain-bb": "FP64 multicolour GS V(2,2) V-cycle, black-box transfers, exact 16^3 solve, in CG",
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


def host_info():
    """host CPU of this box: the threads the oracle may use (the box's share, OMP_NUM_THREADS),
    nproc and the processor model"""
    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS") or nproc)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": threads, "nproc": nproc, "cpu": model}


def cpu_time(h, prm, b, threads, runs):
    """median wall time of `runs` oracle solves on `threads` host threads, after one warm-up"""
    from oracle import pyoracle as po
    po.set_threads(threads)
    h.solve(prm, b)
    ts, steps = [], 0
    for _ in range(runs):
        t0 = time.perf_counter()
        st, x, steps, dn = h.solve(prm, b)
        ts.append(time.perf_counter() - t0)
    po.set_threads(0)
    return float(np.median(ts)), steps, st


def cpu_baselines(threads, deploy, gpu_steps, log=lambda *a: None):
    """The oracle (the C restatement of femlisp's algorithm: FP64, femlisp's pattern and
    iteration semantics, OpenMP over rows and over one colour's rows or blocks) timed on
    `threads` host threads and on one, median of 5 after a warm-up, setup excluded:

      2D Q1 32^2 Poisson, V(2,2) <jacobi> 0.6 PCG to 1e-10: the full solve;
      3D Q1 64^3 Poisson, random rhs, V(2,2) multicolour GS PCG to 1e-10: the full solve;
      3D Q2 48^3, kappa = exp(N(0, 1.5^2)), <geometric-ssc> V(1,1) PCG: runs of 1 and 3
        iterations (1 and 2 on one thread), extrapolated to the GPU's iteration count;
      3D Q1 128^3 inclusions (the deploy workload), femlisp's plain GMG-CG: runs of 2 and 8
        iterations (1 and 3 on one thread), extrapolated to the plain path's iteration count.
    """
    from oracle import pyoracle as po
    out = {}

    def sample(name, h, prm, b, ks_all, ks_one, scale_to, runs_one=3):
        """t(k) = t0 + k c for k iterations (t0: the oracle's per-solve set-up, orderings
        and the coarse factorisation); two sample lengths give c and t0, and the estimate
        of the whole solve is t0 + c x the GPU's iteration count"""
        res = {}
        for tag, th, ks, runs in (("all", threads, ks_all, 5), ("one", 1, ks_one, runs_one)):
            t = []
            for k in ks:
                prm.maxsteps = k - 1
                tk, steps, st = cpu_time(h, prm, b, th, runs)
                assert steps == k
                t.append(tk)
            c = (t[1] - t[0]) / (ks[1] - ks[0])
            t0 = max(t[0] - ks[0] * c, 0.0)
            res[tag] = {"s": t0 + c * scale_to, "per_iteration_s": c, "setup_s": t0}
        log(f"[bench] cpu {name}: {res['all']['s']:.2f}s on {threads} threads, "
            f"{res['one']['s']:.2f}s on one")
        out[name] = {"s": res["all"]["s"], "s_1core": res["one"]["s"], "threads": threads,
                     "iterations": scale_to,
                     "per_iteration_s": res["all"]["per_iteration_s"],
                     "per_iteration_s_1core": res["one"]["per_iteration_s"],
                     "sample": f"runs of {ks_all[0]} and {ks_all[1]} iterations on {threads} "
                               f"threads (median of 5), {ks_one[0]} and {ks_one[1]} on one "
                               f"(median of {runs_one}): per-solve set-up + per-iteration "
                               f"time x {scale_to} iterations"}

    # 2D Q1 32^2 and 3D Q1 64^3: whole solves
    for name, dim, lv, sm, om, sym, rhs in (("2d_q1_32", 2, 6, po.SMOOTH_JACOBI, 0.6, 0, None),
                                           ("3d_q1_64", 3, 7, po.SMOOTH_SOR, 1.0, 1, "random")):
        h = po.Hierarchy(dim, 1, 1, lv)
        b = h.rhs()
        if rhs == "random":
            b = np.random.default_rng(2024).uniform(-1, 1, h.ndofs())
            b[po.boundary(dim, 1, 1 << (lv - 1))] = 0
        prm = po.params(outer=po.OUTER_CG, smoother=sm, omega=om, ordering=po.ORDER_COLOUR,
                        pre_steps=2, post_steps=2, base_level=0, symmetric=sym, tol=1e-10,
                        maxsteps=200)
        ta, steps, st = cpu_time(h, prm, b, threads, 5)
        t1, _, _ = cpu_time(h, prm, b, 1, 5)
        assert st == 0
        out[name] = {"s": ta, "s_1core": t1, "threads": threads, "iterations": steps,
                     "sample": "the whole solve, median of 5"}
        log(f"[bench] cpu {name}: {ta:.3f}s on {threads} threads, {t1:.3f}s on one")
        del h
    # 3D Q2 48^3 with the block SSC femlisp selects for order 2
    if gpu_steps.get("q2") is not None:
        kap = np.exp(np.random.default_rng(48).normal(scale=1.5, size=(48,) * 3))
        t0 = time.perf_counter()
        h = po.Hierarchy(3, 2, 3, 5, kappa=kap)
        su = time.perf_counter() - t0
        log(f"[bench] cpu 3d_q2_48_ssc: oracle setup {su:.1f}s")
        prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SSC, omega=1.0, pre_steps=1,
                        post_steps=1, base_level=0, symmetric=1, tol=1e-10)
        sample("3d_q2_48_ssc", h, prm, h.rhs(), (1, 3), (1, 2), gpu_steps["q2"], runs_one=1)
        out["3d_q2_48_ssc"]["setup_s"] = su
        del h
    # the deploy workload
    kappa, n0, levels, base, b = deploy
    t0 = time.perf_counter()
    h = po.Hierarchy(3, 1, n0, levels, kappa=kappa)
    su = time.perf_counter() - t0
    log(f"[bench] cpu 3d_q1_128_inclusions: oracle setup {su:.1f}s")
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, omega=1.0,
                    ordering=po.ORDER_COLOUR, pre_steps=2, post_steps=2, base_level=base,
                    symmetric=1, tol=1e-10)
    sample("3d_q1_128_inclusions", h, prm, b, (2, 8), (1, 3), gpu_steps["plain"], runs_one=3)
    out["3d_q1_128_inclusions"]["setup_s"] = su
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="cells per axis (power of two)")
    ap.add_argument("--base", type=int, default=3,
                    help="coarse-solve level (2^base cells/axis) of the geometric paths (plain, "
                         "mixed: femlisp's GMG, compared with the oracle at the same level)")
    ap.add_argument("--bb-base", type=int, default=4,
                    help="coarse-solve level of the black-box paths (-bb, -bb-jac, learned): the "
                         "16^3 Galerkin level solved exactly (a 3375 x 3375 inverse)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--train-problems", type=int, default=2048,
                    help="random-coefficient 32^3 problems of the training set (one residual/"
                         "error pair each)")
    ap.add_argument("--train-epochs", type=int, default=2)
    ap.add_argument("--learned-level", type=int, default=32,
                    help="cells per axis of the level the network runs on")
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
    ap.add_argument("--assembly-only", action="store_true",
                    help="only the MFMA work (counter runs): 3D Q1 assembly at n^3 with FP64 and "
                         "BF16 MFMA contractions, 3D Q2 48^3 assembly, 8 U-Net training steps at "
                         "32^3 with batch 32")
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
                    # the value planes the kernels stream: 14 of 27 when the stencil is
                    # exactly symmetric (the lower ones come from the neighbours' upper planes)
                    pl = M.stencil_planes()
                    assert pl in (14, 27)
                    sb, wb = pl * N * vb + 2 * N * vb, pl * N * vb + 5 * N * vb
                    jb = pl * N * vb + 4 * N * vb
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

    def assembly_entry(reps=5):
        """Assembly of the value path's fine matrix: 3D Q1 n^3, FP32 values, BF16-MFMA element
        contractions (k_asm_q1: element matrices formed in LDS, one pass).  Timed: whole
        re-assemblies for the coefficient (values, load of the per-cell factors, refresh of the
        stored forms), HIP events over `reps` calls.  Algorithmic flops (the reference's per-qp
        chain, ellsys-fe.lisp:120-127): cells x nq x 2 (nloc d^2 + nloc^2 d); bytes: the
        coefficient read and the CSR values written once."""
        A, _ = lat.assemble(3, 1, n, (kv, n), prec=F.F32, mode=F.ASM_MFMA, rhs=False)
        lat.reassemble(A, (kv, n), mode=F.ASM_MFMA)
        ctx.timer_start()
        for _ in range(reps):
            lat.reassemble(A, (kv, n), mode=F.ASM_MFMA)
        ms = ctx.timer_stop() / reps
        cells, nq, nloc, d = n ** 3, 8, 8, 3
        flops = cells * nq * 2 * (nloc * d * d + nloc * nloc * d)
        byt = cells * 8 + A.nnz * 4
        out = {"kernel": "k_asm_q1<float, 2> (BF16-MFMA element contractions in LDS) + refresh",
               "ms": ms, "flops": flops, "bytes": byt,
               "tflops": flops / (ms * 1e-3) / 1e12, "gbs": byt / (ms * 1e-3) / 1e9}
        out["frac_hbm"] = out["gbs"] / 8000.0
        del A
        return out

    if args.assembly_only:
        for prec in (F.F64, F.F32):
            A, _ = lat.assemble(3, 1, n, (kv, n), prec=prec, mode=F.ASM_MFMA, rhs=False)
            lat.reassemble(A, (kv, n), mode=F.ASM_MFMA)
            del A
        kq = np.exp(np.random.default_rng(48).normal(scale=1.5, size=(48,) * 3))
        for prec in (F.F64, F.F32):
            A, _ = lat.assemble(3, 2, 48, kq, prec=prec, mode=F.ASM_MFMA, rhs=False)
            del A
        net = learned.UNet(2, 3)
        xb = GpuVector.from_array(np.random.default_rng(1).normal(size=32 * 32 ** 3 * 2), F.F32)
        tb = GpuVector.from_array(np.random.default_rng(2).normal(size=32 * 32 ** 3), F.F32)
        for _ in range(8):
            net.train_step(xb, tb, 32, 32, 1e-3)
        ctx.synchronize()
        log("[bench] assembly-only done")
        return

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
        if name not in paths and not (name == "mixed-bb-jac" and "learned" in paths):
            continue
        t = time.perf_counter()
        base = args.bb_base if tr == "blackbox" else args.base
        if sm == "jac":
            cs = it.GeometricCS(it.Jacobi(0.9), 2, 2, base, prec=prec, transfer=tr)
        else:
            cs = it.GeometricCS(it.SOR(1.0), 2, 2, base, symmetric=True, prec=prec,
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
            # the training configuration: 32^3 cells, kappa = exp(g), g a smoothed Gaussian
            # field (correlation length 4 cells, sigma 2), pairs from the FP64 multigrid,
            # Adam lr 1e-3, batch 32
            X, T, cnt = learned.generate_pairs(n=32, problems=args.train_problems,
                                               snapshots=(1,), kinds=("gaussian",), seed=11)
            t_data = time.perf_counter() - t
            hist, vloss = learned.train(net, X, T, cnt, 32, epochs=args.train_epochs, batch=32,
                                        lr=1e-3, seed=2)
            train_info = {"pairs": cnt, "data_s": t_data,
                          "train_s": time.perf_counter() - t - t_data,
                          "loss": hist, "holdout": vloss}
            del X, T
        # on the value path's hierarchy, at the level with learned_level cells per axis (the
        # coefficient there: the geometric mean of the fine cells')
        net_kappa = learned.attach_coarse(hier["mixed-bb-jac"][1], net, kappa, args.learned_level)

    def run(path):
        learned_p = path == "learned"
        cs, H = hier["mixed-bb-jac" if learned_p else path]
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
    asm = assembly_entry()
    log(f"[bench] assembly 3D Q1 {n}^3 FP32 (BF16 MFMA), whole reassembly: {asm['ms']:.2f} ms, "
        f"{asm['tflops']:.2f} TFLOP/s, {asm['gbs']:.0f} GB/s of the value write")
    kern = fine_kernels(args.kernel_reps)
    for k, v in kern.items():
        log(f"[bench] {k}: {v['ms'] * 1e3:.1f} us, {v['gbs']:.0f} GB/s "
            f"({v['gbs'] / 8000 * 100:.1f}% of 8 TB/s)")
    # the byte models count each value once; a rate above the HBM peak means a model counts
    # too few bytes (or the data came from the Infinity Cache), so it is flagged in the line
    over_peak = sorted(k for k, v in kern.items() if v["gbs"] > 8000.0)
    if over_peak:
        log(f"[bench] WARNING: above the 8 TB/s HBM peak by the byte model: {over_peak}")

    q2 = None if args.no_q2 else q2_config()

    # the first two reference configurations on the GPU, with the oracle's algorithm: 2D Q1
    # 32^2 Poisson, V(2,2) <jacobi> 0.6 PCG, and 3D Q1 64^3 Poisson with a random rhs, V(2,2)
    # multicolour GS PCG, both FP64 to 1e-10 (median of 5 solves after a warm-up)
    small = {}
    for name, dim, lv, sm, sym, rhs in (("2d_q1_32", 2, 6, it.Jacobi(0.6), False, None),
                                       ("3d_q1_64", 3, 7, it.SOR(1.0), True, "random")):
        cs = it.GeometricCS(sm, 2, 2, 0, symmetric=sym)
        Hs = it.Hierarchy.lattice(cs, dim, 1, 1, lv)
        bs = Hs.rhs
        if rhs == "random":
            from oracle import pyoracle as po
            bh = np.random.default_rng(2024).uniform(-1, 1, Hs.top_matrix().nrows)
            bh[po.boundary(dim, 1, 1 << (lv - 1))] = 0
            bs = GpuVector.from_array(bh)
        ts = []
        for k in range(6):
            r = it.solve(it.LinearSolver(it.CG(cs), failure_if=(">", "step", 200)),
                         {"matrix": Hs.top_matrix(), "rhs": bs, "hierarchy": Hs})
            assert r["status"] == "success"
            if k:
                ts.append(r["seconds"])
        small[name] = {"s": float(np.median(ts)), "steps": r["steps"]}
        log(f"[bench] {name}: GPU {small[name]['s'] * 1e3:.2f} ms, {r['steps']} iterations")
        del Hs, bs

    cpu, host = None, host_info()
    if not args.no_cpu:
        log(f"[bench] cpu baselines (oracle on {host['threads']} threads and on one, "
            f"{host['cpu']}, nproc {host['nproc']}) ...")
        cpu = cpu_baselines(host["threads"], (kappa, n0, levels, args.base, b_host),
                            {"plain": results["plain"]["steps"],
                             "q2": q2["plain-ssc"]["steps"] if q2 else None}, log)
        for k, v in cpu.items():
            log(f"[bench] cpu {k}: {v['s']:.3f}s on {v['threads']} threads, "
                f"{v['s_1core']:.3f}s on one ({v['sample']})")
        gpu_s = {"2d_q1_32": small["2d_q1_32"]["s"], "3d_q1_64": small["3d_q1_64"]["s"],
                 "3d_q1_128_inclusions": results["plain"]["s"]}
        if q2:
            gpu_s["3d_q2_48_ssc"] = q2["plain-ssc"]["s"]
        for k in cpu:
            cpu[k]["gpu_plain_s"] = gpu_s.get(k)

    # the metric's path: the learned correction with the mixed-precision cycle; without it (a
    # run without the learned path), the fastest path, named in config
    fast = "learned" if "learned" in results else min(
        (p for p in results if p != "plain"), key=lambda p: results[p]["s"])
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
                               f"30% inclusions r 3-8 cells, base level {args.base} "
                               f"(black-box paths {args.bb_base}); value = "
                               f"'{fast}' path ({PATH_DESC.get(fast, fast)}); plain = femlisp's "
                               f"FP64 GMG-CG, V(2,2) multicolour GS",
                   "global_batch": 1, "seq_len": N, "parallelism": "single GPU"},
        "paths": {p: {k: v for k, v in r.items()} for p, r in results.items()},
        "speedup_vs_plain": {p: results["plain"]["s"] / results[p]["s"] for p in results},
        "learned_training": train_info,
        "learned_gain": None if not ("learned" in results and "mixed-bb-jac" in results) else {
            "steps_without": results["mixed-bb-jac"]["steps"],
            "steps_with": results["learned"]["steps"],
            "time_ratio_without_over_with": results["mixed-bb-jac"]["s"] / results["learned"]["s"],
            "fell_back": results["learned"]["fell_back"]},
        "assembly": asm,
        "setup_s": setup,
        "vcycle_gdofs": {k: v["gdofs"] for k, v in vcyc.items()},
        "vcycle_ms": {k: v["ms"] for k, v in vcyc.items()},
        "kernels": {k: {"us": v["ms"] * 1e3, "GBps": v["gbs"], "frac": v["gbs"] / 8000.0}
                    for k, v in kern.items()},
        "kernels_above_hbm_peak": over_peak,
        "q2_48": q2,
        "roofline": roof if roof is not None else {
            "bound": "hbm", "kernel": "k_q1_sor<float> (fine-level <sor> sweep, per-cell form)",
            "achieved": dom["gbs"], "peak": 8000.0, "unit": "GB/s",
            "frac": dom["gbs"] / 8000.0, "traffic": None},
        "cpu_baseline": None if cpu is None else {
            "value": cpu["3d_q1_128_inclusions"]["s"], "unit": "s", "cores": host["threads"],
            "kind": "port",
            "sample": f"the oracle's plain FP64 GMG-CG on the same {n}^3 system (OpenMP over "
                      f"rows and colours): {cpu['3d_q1_128_inclusions']['sample']} of the plain "
                      f"path; setup excluded",
            "cores_1": cpu["3d_q1_128_inclusions"]["s_1core"],
            "nproc": host["nproc"], "cpu": host["cpu"],
            "configs": cpu},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
