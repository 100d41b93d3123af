#!/usr/bin/env python3
"""Benchmark: numeric factor + solve of the 3D upwind convection-diffusion
7-point stencil on a 64^3 grid (n = 262,144; diffusivity 1, velocity
(1, 0.5, 0.25), cell Peclet 0.5), LU through SparseSolver's numeric phase on
one MI355X, in four precision paths (a fifth on request):

  fp64       factor and solve in double (the reference's Float64 path)
  fp32       factor and solve in float, residual and update in double,
             refined until the normwise backward error is <= 1e-14
  fp32_bf16  as fp32, with bf16 operands (float accumulation) in the
             trailing updates of fronts wider than 256 columns
  fp32_bf16x3  as fp32_bf16 with split-bf16 operands (x = hi + lo, three bf16
             products per pair, about 16 significant bits)
  fp32_fp16  as fp32_bf16 with fp16 operands under diagonal scaling (not in
             the default set: at 64^3 it needs as many refinement steps as
             fp32_bf16 and its scaled operand staging costs more)

A step is: inmatrix (values already resident in HBM) -> factor -> solve of
one right-hand side b = A x_true (x_true ~ U[-1, 1], splitmix64 seed 2) with
the backward-error certificate (and refinement for fp32). The host symbolic
phase (MMD, etree, supernodes) runs once before the timed region, as in
sparspaklu! refactorisation, and so does the autotuner that picks each path's
plan knobs (front split, block width, update depth) by timing its factor
(--no-autotune keeps the defaults).

Beside the headline, the same steps run on the mixed-precision
configuration: Q1 linear elasticity on a 40^3 element mesh (3 dof per node,
n = 206,763, z = 0 face clamped as identity rows, body load N(0, 1)), in
fp64, fp32_bf16x3 and fp32_bf16 (key "elasticity" of the JSON line, with the
best mixed path's factor+solve speed-up over fp64). Prints one JSON line.

--batched adds the two batched configurations (key "batched"): 2 000
independent 2D 5-point Laplacians (sides uniform in 8..96, splitmix64 seed 4;
a diagonal shift U[0.1, 1] each, seed 40; SPD, 4 right-hand sides each) in
fp64 and fp32 + refinement, and 256 value sets of one 32^3 convection-
diffusion pattern (log-uniform diffusivity in [1e-2, 1e2] per cell, seeds
5000 + k; 16 right-hand sides each) in fp32 and fp32_bf16x3 + refinement.
Each path is autotuned, then timed over the same steps (values and
right-hand sides resident in HBM; factor, then the solve and refinement of
every right-hand side of every system). Not in the default run (the systems
take about a minute to generate on the host).

    python bench.py [--steps K] [--warmup W] [--grid 64] [--cpu-grid 36] [--elastic 40]
                    [--batched]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "sparspak.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

# MI355X dense matrix TFLOP/s (datasheet) of the arithmetic each kernel class
# runs in: everything in double on the fp64 path, in float on the fp32 path;
# on the fp32_bf16 (fp32_fp16) path the trailing updates of the large fronts
# run on the bf16 (fp16) matrix cores, same dense peak, and the rest in float
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2516.6}
HBM_PEAK = 8000.0                       # GB/s


def kernel_peak(path, kclass):
    if path == "fp64":
        return PEAK["fp64"]
    if path in ("fp32_bf16", "fp32_fp16", "fp32_bf16x3") and kclass == "large_update":
        return PEAK["bf16"]
    return PEAK["fp32"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--cpu-grid", type=int, default=36)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--paths", default="fp64,fp32,fp32_bf16,fp32_bf16x3")
    ap.add_argument("--elastic", type=int, default=40,
                    help="elements per side of the elasticity configuration measured "
                         "beside the headline (0: skip)")
    ap.add_argument("--elastic-paths", default="fp64,fp32_bf16x3,fp32_bf16")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default plan knobs instead of autotuning each path")
    ap.add_argument("--tuning", default=None,
                    help="small_lmax,block_cols,upd_group,small_wave_min for every path "
                         "instead of autotuning (e.g. to profile the plan a bench run chose)")
    ap.add_argument("--batched", action="store_true",
                    help="also run the two batched configurations (key \"batched\")")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (three times), the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU path (subtree-to-GPU "
              "mapping with one update exchange over RCCL) is a design, not built; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.calibrate_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {src.numel() * 8} bytes three times")
        return
    from sparspak_hip.gallery import convdiff3d, elasticity3d, normal, uniform

    g = args.grid
    t0 = time.time()
    A = convdiff3d(g)
    n = A.shape[0]
    xtrue = uniform(2, n, -1.0, 1.0)
    b = A @ xtrue
    log(f"matrix convdiff3d {g}^3 n={n} nnz={A.nnz} ({time.time() - t0:.1f}s)")
    results, F, t_sym = run_workload(A, b, xtrue, args.paths.split(","), args)

    # the mixed-precision configuration: Q1 elasticity on an ne^3 mesh,
    # 3 dof per node, z = 0 face clamped (identity rows kept), body load
    # N(0, 1) per free dof (splitmix64 seed 3), 0 on the clamped dofs
    elastic = None
    if args.elastic > 0:
        ne = args.elastic
        t0 = time.time()
        Ae = elasticity3d(ne)
        ne_n = Ae.shape[0]
        be = normal(3, ne_n)
        be[:3 * (ne + 1) ** 2] = 0.0
        log(f"matrix elasticity3d {ne}^3 n={ne_n} nnz={Ae.nnz} ({time.time() - t0:.1f}s)")
        eres, eF, et_sym = run_workload(Ae, be, None, args.elastic_paths.split(","), args)
        elastic = {
            "workload": f"Q1 linear elasticity {ne}^3 elements, 3 dof/node, E=1, nu=0.3, "
                        "z=0 clamped (identity rows), body load N(0,1); LU numeric phase",
            "n": ne_n, "nnz": int(Ae.nnz), "F_LU": eF, "host_symbolic_s": et_sym,
            "paths": {p: {k: v for k, v in eres[p].items() if k != "kernels"} for p in eres},
            "kernels": {p: eres[p]["kernels"] for p in eres},
        }
        if "fp64" in eres:
            ok = [p for p in eres if p != "fp64" and eres[p]["backward_error"] <= 1e-14]
            if ok:
                best = min(ok, key=lambda p: eres[p]["factor_solve_s"])
                elastic["best_mixed_path"] = best
                elastic["speedup_vs_fp64"] = (eres["fp64"]["factor_solve_s"]
                                              / eres[best]["factor_solve_s"])
                log(f"elasticity: {best} factor+solve {eres[best]['factor_solve_s'] * 1e3:.1f} ms"
                    f" vs fp64 {eres['fp64']['factor_solve_s'] * 1e3:.1f} ms "
                    f"({elastic['speedup_vs_fp64']:.2f}x)")

    # headline: the fastest path that certifies eta <= 1e-14
    ok = [p for p in results if results[p]["backward_error"] <= 1e-14]
    head = min(ok or list(results), key=lambda p: results[p]["factor_solve_s"])
    R = results[head]
    kst = R["kernels"]
    dom = max(kst, key=lambda k: kst[k]["ms"])
    dk = kst[dom]
    compute_bound = dom in ("large_update", "small_front", "large_diag", "large_offdiag",
                            "large_diag_border")
    prec_name = head
    # a compute kernel is MFMA-bound only when its algorithmic intensity
    # (flops per algorithmic HBM byte) times the HBM peak exceeds the matrix
    # peak of its arithmetic; otherwise its roof is HBM
    intensity = dk["flops"] / dk["bytes"] if dk["bytes"] > 0 else float("inf")
    mfma_roof = compute_bound and dk["flops"] > 0 and \
        kernel_peak(head, dom) <= intensity * HBM_PEAK / 1e3
    if mfma_roof:
        achieved = dk["flops"] / (dk["ms"] / 1e3) / 1e12
        peak = kernel_peak(head, dom)
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                "kernel": dom, "launches": dk["launches"],
                "avg_launch_ms": dk["ms"] / dk["launches"],
                "work_per_launch": dk["flops"] / dk["launches"]}
    else:
        achieved = dk["bytes"] / (dk["ms"] / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": None, "kernel": dom,
                "launches": dk["launches"], "avg_launch_ms": dk["ms"] / dk["launches"],
                "work_per_launch": dk["bytes"] / dk["launches"]}
        if dk["flops"] > 0:
            roof["flops_per_byte"] = intensity
            roof["achieved_tflops"] = dk["flops"] / (dk["ms"] / 1e3) / 1e12
            roof["matrix_peak_tflops"] = kernel_peak(head, dom)

    # measured HBM bytes per launch of the dominant kernel, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the same workload (scaled
    # by a plain copy's counter reading); None when no summary covers it
    tr = traffic_from_profile(roof["kernel"], prec_name)
    if tr is not None:
        roof["traffic"] = tr
        roof["traffic_source"] = "profiles/" + TRAFFIC_FILE

    batched = run_batched(args) if args.batched else None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_grid)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "factor+solve s and factor TFLOP/s: FP32+IR vs FP64 at <=1e-14 backward err",
        "value": R["factor_tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": R["factor_solve_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32+f64ir", "fp32_bf16": "f32+bf16upd+f64ir",
                  "fp32_fp16": "f32+fp16upd+f64ir", "fp32_bf16x3": "f32+bf16x3upd+f64ir"}[head],
        "data": "synthetic (generated matrix and rhs, splitmix64 seeds)",
        "config": {"workload": f"3D upwind convection-diffusion 7-point {g}^3 LU "
                               "(SparseSolver numeric phase: inmatrix+factor+solve)",
                   "n": n, "nnz": int(A.nnz), "F_LU": F, "headline_path": head,
                   "host_symbolic_s": t_sym, "plan_knobs": R["tuning"] or "default"},
        "paths": {p: {k: v for k, v in results[p].items() if k != "kernels"} for p in results},
        "kernels": {p: results[p]["kernels"] for p in results},
        "roofline": roof,
        "cpu_baseline": cpu,
        "elasticity": elastic,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


def run_batched(args):
    """The batched configurations through BatchedSolver (one device factor
    for all systems): per path autotune, warmup, then the timed steps
    (factor; solve of every right-hand side, refined for the float paths),
    times from the handle's HIP events, each system's backward error
    checked on the last step."""
    import torch
    from sparspak_hip.batch import BatchedSolver
    from sparspak_hip.gallery import convdiff3d_varcoef, laplace2d, splitmix64, uniform

    out = {}
    configs = []
    nsys = 2000
    sides = 8 + (splitmix64(4, nsys) % np.uint64(89)).astype(int)
    shifts = uniform(40, nsys, 0.1, 1.0)
    configs.append(("laplace2d_uneven_2000_spd", "spd", lambda: [
        laplace2d(int(sd), shift=float(sh)) for sd, sh in zip(sides, shifts)], 4, ["fp64", "fp32"],
        "2 000 2D 5-point Laplacians, sides U{8..96}, shift U[0.1,1], SPD, 4 rhs each"))
    g = 32
    configs.append(("convdiff3d_32_value_sets_256", "lu", lambda: [
        convdiff3d_varcoef(g, 10.0 ** uniform(5000 + k, g ** 3, -2.0, 2.0)) for k in range(256)],
        16, ["fp32", "fp32_bf16x3"],
        "256 value sets of one 32^3 convection-diffusion pattern, kappa log-U[1e-2,1e2], "
        "velocity (1,0,0), 16 rhs each"))
    for key, kind, gen, nrhs, paths, desc in configs:
        t0 = time.time()
        mats = gen()
        bs = BatchedSolver(mats, kind=kind)
        bs.set_values()
        n = bs.n
        rhs = np.stack([uniform(900 + k, n, -1.0, 1.0) for k in range(nrhs)], axis=1)
        d_b = torch.from_numpy(np.ascontiguousarray(rhs.T)).cuda()  # column-major n x nrhs
        d_x = torch.empty_like(d_b)
        log(f"batched {key}: n={n} systems={len(mats)} value_sets={bs.value_sets} "
            f"({time.time() - t0:.1f}s)")
        res = {"workload": desc, "n": n, "systems": len(mats), "value_sets": bool(bs.value_sets),
               "nrhs_per_system": nrhs, "F": bs.flops(), "paths": {}}
        for path in paths:
            refine = path != "fp64"
            tuning = bs.autotune(path)

            def step():
                bs.factor(path)
                d_x.copy_(d_b)
                bs.solve_device(d_x, nrhs, refine=refine)
                st = bs.stats()
                return st.t_factor_ms, (st.t_refine_ms if refine else st.t_solve_ms), st

            for _ in range(args.warmup):
                step()
            torch.cuda.synchronize()
            tf, ts = [], []
            for _ in range(args.steps):
                f_ms, s_ms, st = step()
                tf.append(f_ms)
                ts.append(s_ms)
            torch.cuda.synchronize()
            x = d_x.cpu().numpy().T
            worst = 0.0
            off = 0
            for A in mats:
                m = A.shape[0]
                xs, bsys = x[off:off + m], rhs[off:off + m]
                r = bsys - A @ xs
                an = abs(A).sum(axis=1).max()
                eta = np.max(np.abs(r), axis=0) / (an * np.max(np.abs(xs), axis=0)
                                                   + np.max(np.abs(bsys), axis=0))
                worst = max(worst, float(np.max(eta)))
                off += m
            f_ms, s_ms = float(np.median(tf)), float(np.median(ts))
            res["paths"][path] = {"factor_s": f_ms / 1e3, "solve_refine_s": s_ms / 1e3,
                                  "factor_solve_s": (f_ms + s_ms) / 1e3,
                                  "factor_tflops": bs.flops() / (f_ms / 1e3) / 1e12,
                                  "refine_iters": int(st.refine_iters),
                                  "worst_backward_error": worst, "tuning": tuning}
            log(f"batched {key} {path}: factor {f_ms:.1f} ms "
                f"({bs.flops() / f_ms / 1e9:.2f} TF/s), solve{'+ir' if refine else ''} "
                f"{s_ms:.1f} ms, iters {st.refine_iters}, worst eta {worst:.2e}")
        out[key] = res
        del bs
    return out


def run_workload(A, b, xtrue, paths, args):
    """Symbolic phase and upload once, then per precision path: autotune
    (outside the timed steps), warmup, the timed steps (inmatrix from
    device-resident values -> factor -> refined solve of b), and one extra
    step under the per-launch profile. Returns (results per path, F_LU,
    host symbolic seconds)."""
    import torch
    from sparspak_hip import SparseSolver, findorder, symbolicfactor
    from sparspak_hip import _lib

    dev = torch.device("cuda:0")
    n = A.shape[0]
    t0 = time.time()
    s = SparseSolver(A)
    findorder(s)
    symbolicfactor(s)
    t_sym = time.time() - t0
    sym = s.slvr.sym
    F = sym.flops()
    log(f"host symbolic {t_sym:.2f}s nsuper={sym.nsuper} nnz(L)={sym.nnzl} F_LU={F:.4e}")
    t0 = time.time()
    s.slvr._upload()
    dv = s.slvr._dev
    L = dv.L
    h = dv.h
    log(f"upload {time.time() - t0:.2f}s")

    d_nz = torch.from_numpy(np.ascontiguousarray(A.data, dtype=np.float64)).to(dev)
    d_b = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    d_x = torch.empty_like(d_b)
    # the library orders its reads of d_nz / d_x after torch's stream
    _lib.check(L.spkhip_set_caller_stream(
        h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "set_caller_stream")
    torch.cuda.synchronize()
    anorm = abs(A).sum(axis=1).max()

    prec = {"fp64": 0, "fp32": 1, "fp32_bf16": 2, "fp32_fp16": 3, "fp32_bf16x3": 4}

    def step(path):
        _lib.check(L.spkhip_inmatrix_device(h, ctypes.c_void_p(d_nz.data_ptr())), "inmatrix")
        _lib.check(L.spkhip_factor(h, prec[path]), "factor")
        d_x.copy_(d_b)
        rc = L.spkhip_solve_refine_device(h, 1, ctypes.c_void_p(d_x.data_ptr()), n, 1e-14, 10)
        if rc not in (0, -5):
            _lib.check(rc, "solve")
        st = _lib.Stats()
        L.spkhip_stats(h, ctypes.byref(st))
        return st

    results = {}
    for path in paths:
        tuning, t_tune = None, 0.0
        if args.tuning:
            vals = [int(v) for v in args.tuning.split(",")]
            tu = _lib.Tuning(*vals, 0.0)
            _lib.check(L.spkhip_set_tuning(h, ctypes.byref(tu)), "set_tuning")
            tuning = dict(zip(_lib.TUNING_KNOBS, vals))
        elif not args.no_autotune:
            # plan knobs chosen by timing this path's factorisation (setup,
            # like the symbolic phase; outside the timed steps)
            t0 = time.perf_counter()
            _lib.check(L.spkhip_inmatrix_device(h, ctypes.c_void_p(d_nz.data_ptr())), "inmatrix")
            tu = _lib.Tuning()
            _lib.check(L.spkhip_autotune(h, prec[path], 2, ctypes.byref(tu)), "autotune")
            t_tune = time.perf_counter() - t0
            tuning = {f: getattr(tu, f) for f in _lib.TUNING_KNOBS}
            log(f"{path}: autotune {t_tune:.1f}s -> {tuning} ({tu.factor_ms:.1f} ms)")
        for _ in range(args.warmup):
            step(path)
        torch.cuda.synchronize()
        tf, ts = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = step(path)
            tf.append(st.t_factor_ms)
            ts.append(st.t_refine_ms)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        x = d_x.cpu().numpy()
        r = b - A @ x
        eta = float(np.max(np.abs(r)) / (anorm * np.max(np.abs(x)) + np.max(np.abs(b))))
        fwd = (None if xtrue is None else
               float(np.max(np.abs(x - xtrue)) / np.max(np.abs(xtrue))))
        # per-kernel profile on one extra, untimed step
        L.spkhip_profile_enable(h, 1)
        step(path)
        kst = {}
        for k, name in enumerate(_lib.KCLASSES):
            ks = _lib.KStat()
            L.spkhip_profile_get(h, k, ctypes.byref(ks))
            if ks.launches:
                kst[name] = {"launches": ks.launches, "ms": ks.ms, "flops": ks.flops,
                             "bytes": ks.bytes}
        L.spkhip_profile_enable(h, 0)
        t_factor = float(np.median(tf)) / 1e3
        t_solve = float(np.median(ts)) / 1e3
        results[path] = {
            "factor_s": t_factor, "solve_refine_s": t_solve,
            "factor_solve_s": t_factor + t_solve, "step_wall_s": wall,
            "factor_tflops": F / t_factor / 1e12,
            "refine_iters": int(st.refine_iters), "backward_error": eta,
            "forward_error_vs_xtrue": fwd, "kernels": kst,
            "tuning": tuning, "autotune_s": t_tune,
        }
        log(f"{path}: factor {t_factor * 1e3:.1f} ms ({F / t_factor / 1e12:.2f} TF/s), "
            f"solve+ir {t_solve * 1e3:.1f} ms, iters {st.refine_iters}, eta {eta:.2e}, "
            f"wall/step {wall * 1e3:.1f} ms")
        for k, v in kst.items():
            log(f"   {k:16s} {v['launches']:6d} launches {v['ms']:9.3f} ms "
                f"{v['flops'] / max(v['ms'], 1e-9) / 1e9:9.2f} TF/s "
                f"{v['bytes'] / max(v['ms'], 1e-9) / 1e6:9.1f} GB/s")
    del d_nz, d_b, d_x
    s.slvr._dev = None
    return results, F, t_sym


TRAFFIC_FILE = "hbm_traffic_convdiff3d64.md"


def traffic_from_profile(kernel_class, prec_name):
    """Measured HBM bytes per launch of kernel_class on path prec_name from
    the committed FETCH_SIZE / WRITE_SIZE summary (lines 'traffic <path>
    <class>: <bytes>'), or None when it does not cover them."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    key = f"traffic {prec_name} {kernel_class}"
    with open(path) as f:
        for line in f:
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    return None


def cpu_baseline(grid):
    """The oracle (serial C++ restatement of Sparspak.jl's LU numeric phase,
    generic BLAS-fragment loops) on a bounded sample of the same workload
    family: the convdiff3d grid^3 matrix, factor + solve on one core."""
    import oracle.oracle as orc
    from sparspak_hip.gallery import convdiff3d, uniform
    from sparspak_hip.symbolic import Graph, SymbolicFactor

    A = convdiff3d(grid)
    n = A.shape[0]
    cp = A.indptr.astype(np.int64) + 1
    rv = A.indices.astype(np.int64) + 1
    s = SymbolicFactor(n, Graph.from_csc(n, cp, rv))
    s.findorder()
    s.symbolicfactor()
    F = s.flops()
    t0 = time.perf_counter()
    lnz, unz, ipiv = orc.inmatrix(s, cp, rv, A.data)
    orc.lufactor(s, lnz, unz, ipiv)
    t1 = time.perf_counter()
    b = A @ uniform(2, n, -1.0, 1.0)
    orc.solve(s, lnz, unz, ipiv, b)
    t2 = time.perf_counter()
    return {"value": F / (t1 - t0) / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": f"oracle LU numeric phase on convdiff3d {grid}^3 (n={n}, F_LU={F:.3e}): "
                      f"inmatrix+factor {t1 - t0:.2f} s, solve {t2 - t1:.3f} s",
            "factor_s": t1 - t0, "solve_s": t2 - t1, "nproc": os.cpu_count()}


if __name__ == "__main__":
    main()
