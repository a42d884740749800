"""Benchmark of the nbots FEM solve path on one MI355X.

Workload (default): 2048x1024 cantilever, 4,194,304 T3 elements,
4,200,450 DOF, lognormal(0,1) per-element modulus, clamped left edge,
uniform shear traction on the right edge.  K, F and all solver vectors are
resident in HBM before the timed region.  One step = one FP64
Jacobi-PCG solve (the reference's algorithm) from u = 0 to relative residual
||F - K u|| / ||F|| <= 1e-8.

The JSON line reports DOF * CG-iteration / s over the timed steps (whole
job), the time to 1e-8 of FP64 PCG and of the mixed-precision solver
(FP32 inner work, FP64 residual and update) on the same GPU, the HBM
roofline of the dominant kernel (the FP64 SpMV with its p.w partials, timed
with HIP events on the solver's stream), and the reference's own CPU
Jacobi-PCG timed on this host over a bounded number of iterations.  A SIMP
topology-optimisation loop on a 1024x512 MBB half-beam (1.05M DOF, 50
iterations) gives the ms per SIMP iteration with mixed-precision and with
FP64 solves, and the largest relative difference of their compliance
histories.

    python bench.py [--steps K] [--warmup W] [--config c3|c2|c4] [--no-cpu] [--mixed cheb]
                    [--simp-iters 50] [--no-simp]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from nbots_amd.probe import spmv_bytes, sym_spmv_bytes  # noqa: E402  (algorithmic bytes of one SpMV)

METRIC = ("DOF·CG-iter/s + time-to-1e-8 residual (FP32+IR vs FP64); "
          "ms per SIMP iteration")
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of k_pcg_A measured by two rocprofv3 --pmc passes
# (FETCH_SIZE x 1.99995 + WRITE_SIZE x 1.000, the factors calibrated on a
# 1 GiB elementwise kernel), summarised in profiles/k_pcg_A_c3_hbm_traffic.md;
# keyed by (config, symmetric layout)
PMC_TRAFFIC = {("c3", False): 599187989, ("c3", True): 411256645}
CONFIGS = {
    "c3": dict(nx=2048, ny=1024, name="C3 mixed precision: 2048x1024 cantilever, lognormal(0,1) E, "
                                      "right-edge shear traction"),
    "c2": dict(nx=1024, ny=512, name="C2: 1024x512 cantilever, lognormal(0,1) E, right-edge shear traction"),
    "c4": dict(nx=512, ny=256, name="512x256 cantilever, lognormal(0,1) E (quick)"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pcg_iter_bytes(n_nodes, n_blocks, prec, sym=False):
    """Algorithmic bytes of one CG iteration as implemented: SpMV kernel
    (above) + update kernel (FP64: read g,w,diag, write g = 4 vectors;
    FP32 scaled: read y,p,r,w, write y,r = 6) + direction kernel (FP64: read
    p,g,diag and write the next p = 4 vectors, and on every fourth iteration
    also read the three previous p and read and write x = 9; FP32: read r,p,
    write p = 3)."""
    sv = 8 if prec == 64 else 4
    nv = 9.25 if prec == 64 else 9
    spmv = sym_spmv_bytes(n_nodes, n_blocks) if sym else spmv_bytes(n_nodes, n_blocks, prec)
    return spmv + nv * (2 * n_nodes) * sv


def cpu_baseline(f, p, seconds, threads):
    """The reference's nb_sparse_solve_CG_precond_Jacobi (compiled from its
    own sources into oracle/_ref) on the same K and F, cut after about
    `seconds` of work (10 calibration iterations decide the count), on
    `threads` OpenMP threads and again on 1 (the reference FEM driver's own
    setting, static_elasticity2D.c:88-90) for a quarter of the time; falls
    back to this repo's C restatement of it (1 thread)."""
    import ctypes as C
    import oracle
    rp, col, val, F = f.system()
    kind = "reference"
    try:
        R = oracle.ref()
    except Exception:
        R = None
        kind = "port"
    it, tr = C.c_uint32(0), C.c_double(0)
    if R is not None:
        h = R.ref_sparse_import(p.N, rp, col, val)

        def run(n, nt):
            x = np.zeros(p.N)
            t0 = time.perf_counter()
            R.ref_pcg_handle(h, F, x, n, 0.0, nt, C.byref(it), C.byref(tr))
            return time.perf_counter() - t0
    else:
        threads = 1

        def run(n, nt):
            x = np.zeros(p.N)
            t0 = time.perf_counter()
            oracle.lib().orc_pcg_jacobi(p.N, rp, col, val, F, x, n, 0.0, C.byref(it), C.byref(tr))
            return time.perf_counter() - t0

    def sample(secs, nt):
        t_cal = run(10, nt)
        n = int(max(10, min(20000, secs / max(t_cal / 10, 1e-6))))
        dt = run(n, nt)
        return it.value, dt

    n_it, dt = sample(seconds, threads)
    one = None
    if R is not None and threads > 1:
        n1, dt1 = sample(seconds / 4, 1)
        one = dict(value=p.N * n1 / dt1, iters=n1, seconds=dt1)
    if R is not None:
        R.ref_sparse_free(h)
    return dict(value=p.N * n_it / dt, unit="DOF*CG-iter/s", cores=threads, kind=kind,
                sample=f"{n_it} Jacobi-PCG iterations on the same {p.N}-DOF system "
                       f"(reference build gcc -O2, {threads} OpenMP threads), {dt:.2f} s; "
                       f"time to 1e-8 extrapolated from the FP64 GPU iteration count",
                seconds=dt, iters=n_it, one_thread=one)


def simp_bench(args, opts):
    """1024x512 MBB half-beam (1,051,650 DOF), volume fraction 0.5, penalty 3,
    density filter radius 1.5, OC move 0.2, from rho = 0.5: the SIMP loop
    run with mixed-precision solves and again with FP64 solves (both to
    relative residual 1e-8, warm-started from the previous displacement)."""
    from nbots_amd import synth
    from nbots_amd.fem import Fem, Simp, SOLVE_FP64, SOLVE_MIXED, solve_opts
    nx, ny = 1024, 512
    p = synth.mbb(nx, ny, rho=0.5)
    out = dict(simp_workload=f"{nx}x{ny} MBB half-beam, {p.N} DOF, {p.N_elem} T3, volfrac 0.5, "
                             "penal 3, filter radius 1.5, OC move 0.2, rho0 = 0.5",
               simp_iterations=args.simp_iters)
    runs = {}
    o64 = solve_opts(true_residual=True, simp=True)
    for name, mode, o in (("mixed", SOLVE_MIXED, opts), ("fp64", SOLVE_FP64, o64)):
        f = Fem.from_problem(p)
        t0 = time.perf_counter()
        s = Simp(f, rmin=1.5, volfrac=0.5, penal=3.0, E0=1.0, Emin=1e-9, move=0.2)
        t_setup = time.perf_counter() - t0
        hist, ch, its, solve_ms = [], [], [], []
        t_loop0 = time.perf_counter()
        for it in range(args.simp_iters):
            C, change, info = s.iterate(mode, tol=1e-8, opts=o)
            hist.append(C)
            ch.append(change)
            its.append(info["iters"])
            solve_ms.append(info["ms"])
            if it % 10 == 0 or it == args.simp_iters - 1:
                log(f"[bench] SIMP {name} it {it}: C {C:.6e} change {change:.4f} "
                    f"solve {info['iters']} it {info['ms']:.1f} ms rel_res {info['rel_res']:.1e}")
        t_loop = time.perf_counter() - t_loop0
        runs[name] = dict(hist=np.array(hist), ms=1e3 * t_loop / args.simp_iters,
                          solve_ms=float(np.mean(solve_ms)), iters=its, setup_s=t_setup,
                          change=ch)
        s.destroy()
        f.destroy()
    m, d = runs["mixed"], runs["fp64"]
    rel = np.abs(m["hist"] - d["hist"]) / np.abs(d["hist"])
    out.update(
        ms_per_simp_iter=round(m["ms"], 2),
        ms_per_simp_iter_fp64=round(d["ms"], 2),
        simp_solve_ms_mean=round(m["solve_ms"], 2),
        simp_solve_ms_mean_fp64=round(d["solve_ms"], 2),
        simp_filter_setup_s=round(m["setup_s"], 3),
        simp_solver_iters_mean=round(float(np.mean(m["iters"])), 1),
        simp_solver_iters_mean_fp64=round(float(np.mean(d["iters"])), 1),
        simp_compliance_first_last=[float(d["hist"][0]), float(d["hist"][-1])],
        simp_compliance_hist_max_rel_diff=float(rel.max()),
        simp_change_last=float(m["change"][-1]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--mixed", default="cheb", choices=["cheb", "restart", "reliable"])
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-max-iter", type=int, default=200000)
    ap.add_argument("--cheb-degree", type=int, default=None,
                    help="library default 40")
    ap.add_argument("--cheb-ratio", type=float, default=None,
                    help="library default 3e4")
    ap.add_argument("--precond-bits", type=int, default=16, choices=[16, 32])
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--simp-iters", type=int, default=50)
    ap.add_argument("--no-simp", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X only (asked for --gpus {args.gpus}); "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    from nbots_amd import synth
    from nbots_amd.fem import (Fem, SOLVE_FP64, SOLVE_MIXED, MIXED_CHEB_FCG, MIXED_CG_RESTART,
                               MIXED_CG_RELIABLE, solve_opts)

    cfg = CONFIGS[args.config]
    t0 = time.perf_counter()
    p = synth.cantilever_traction(cfg["nx"], cfg["ny"], seed=args.seed)
    log(f"[bench] {cfg['name']}: {p.N} DOF, {p.N_elem} elements (host gen {time.perf_counter()-t0:.1f}s)")
    t0 = time.perf_counter()
    f = Fem(p.nod, p.conn)
    t_pattern = time.perf_counter() - t0
    t0 = time.perf_counter()
    f.assemble(p.E, p.nu, p.thickness, p.elem_E)
    f.set_bconditions(p)
    t_assemble = time.perf_counter() - t0
    N, nnz, nblk = f.sizes()
    n_nodes = N // 2
    n_valid = nnz // 4
    log(f"[bench] pattern {t_pattern:.2f}s, assembly+BC {t_assemble:.2f}s, nnz {nnz}, "
        f"stored blocks {nblk} (valid {n_valid})")

    # Steps: FP64 Jacobi-PCG solves to 1e-8 (the reference's algorithm,
    # device resident); value = DOF * CG iterations / s over the timed steps.
    # both solvers stop on the true residual ||F - K u|| / ||F|| <= 1e-8
    # (restarting from the iterate when their recursively updated residual
    # has drifted from it)
    o64 = solve_opts(true_residual=True)
    for w in range(args.warmup):
        info = f.solve(SOLVE_FP64, tol=1e-8, max_iter=200000, opts=o64)
        log(f"[bench] warmup {w}: FP64 PCG {info['iters']} it, {info['ms']/1e3:.3f}s")
    times, iters, infos = [], [], []
    t_wall0 = time.perf_counter()
    for s in range(args.steps):
        t0 = time.perf_counter()
        info = f.solve(SOLVE_FP64, tol=1e-8, max_iter=200000, opts=o64)
        dt = time.perf_counter() - t0
        times.append(dt)
        iters.append(info["iters"])
        infos.append(info)
        log(f"[bench] step {s}: FP64 PCG {info['iters']} it, {dt:.3f}s wall, "
            f"{info['ms']/1e3:.3f}s device, rel_res {info['rel_res']:.2e}, "
            f"spmv {info['ms_spmv']*1e3:.1f}us")
    t_wall = time.perf_counter() - t_wall0
    u64 = f.displacement()
    # ||F - K u|| / ||F|| of the solution, evaluated afresh: the solvers stop
    # on their recursively updated residual, as the reference does
    true64 = (f.true_residual(True), f.true_residual(False))
    value = N * sum(iters) / t_wall
    ms_step = 1e3 * t_wall / args.steps

    # dominant kernel: the FP64 SpMV (+ p.w partials), sampled with HIP
    # events on the solver's stream during the timed steps
    spmv_ms = float(np.mean([i["ms_spmv"] for i in infos if i["ms_spmv"] > 0]))
    sym = infos[-1]["layout"] == 1
    b64 = spmv_bytes(n_nodes, n_valid, 64)
    bk = sym_spmv_bytes(n_nodes, n_valid) if sym else b64
    achieved = bk / (spmv_ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=PMC_TRAFFIC.get((args.config, sym)),
                    kernel=("k_pcg_A<OpSym> (FP64 SpMV, symmetric 2x2-block sliced ELL, "
                            "+ p.w partials)" if sym else
                            "k_pcg_A<OpS2> (FP64 2x2-block sliced-ELL SpMV + p.w partials)"),
                    bytes_per_launch=bk, us_per_launch=round(spmv_ms * 1e3, 2))

    # mixed precision, timed separately: time to 1e-8 against the FP64 steps
    extra = {}
    if not args.no_mixed:
        variant = dict(cheb=MIXED_CHEB_FCG, restart=MIXED_CG_RESTART,
                       reliable=MIXED_CG_RELIABLE)[args.mixed]
        opts = solve_opts(variant, cheb_degree=args.cheb_degree, cheb_ratio=args.cheb_ratio,
                          precond_bits=args.precond_bits, true_residual=True)
        f.solve(SOLVE_MIXED, tol=1e-8, max_iter=args.mixed_max_iter, opts=opts)
        mtimes, minfos = [], []
        for s in range(2):
            t0 = time.perf_counter()
            info = f.solve(SOLVE_MIXED, tol=1e-8, max_iter=args.mixed_max_iter, opts=opts)
            mtimes.append(time.perf_counter() - t0)
            minfos.append(info)
            log(f"[bench] mixed ({args.mixed}) {s}: {info['iters']} it, {info['outer']} inner, "
                f"{mtimes[-1]:.3f}s, rel_res {info['rel_res']:.2e}, fallback {info['fallback']}")
        u32 = f.displacement()
        true32 = (f.true_residual(True), f.true_residual(False))
        Fl = f.load_vector()
        c64, c32 = float(Fl @ u64), float(Fl @ u32)
        t_mixed = float(np.median(mtimes))
        t_fp64 = float(np.median(times))
        ms32 = f.spmv_ms(32, 20)
        ms16 = f.spmv_ms(16, 20)
        ms64 = f.spmv_ms(64, 20)
        ms65 = f.spmv_ms(65, 20) if sym else None
        extra = dict(
            time_to_1e8_fp64_s=round(t_fp64, 4),
            time_to_1e8_mixed_s=round(t_mixed, 4),
            mixed_speedup_vs_fp64=round(t_fp64 / t_mixed, 3),
            mixed_variant=args.mixed,
            mixed_solver=("FP64 flexible CG, Chebyshev-Jacobi preconditioner in FP32 arithmetic on "
                          f"{'fp16' if opts.precond_bits == 16 else 'FP32'} matrix values "
                          f"(degree {opts.cheb_degree}, [lmax/{opts.cheb_ratio:g}, lmax])"
                          if variant == MIXED_CHEB_FCG else
                          "FP32 CG sweeps (inner rtol 1e-4) in FP64 iterative refinement"
                          if variant == MIXED_CG_RESTART else
                          "FP32 CG with FP64 reliable updates"),
            mixed_precond_matrix=(("fp16 copy without diagonal blocks"
                                   + (", 16-bit column deltas" if minfos[-1]["layout"] & 4 else ""))
                                  if minfos[-1]["layout"] & 2 else "full scaled copy"),
            mixed_iters=minfos[-1]["iters"], mixed_inner=minfos[-1]["outer"],
            mixed_rel_res=minfos[-1]["rel_res"], mixed_fallback=minfos[-1]["fallback"],
            mixed_true_rel_res=true32[0], mixed_true_rel_res_fp64_eval=true32[1],
            mixed_vs_fp64_disp_rel_l2=float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
            mixed_vs_fp64_compliance_rel=abs(c32 - c64) / abs(c64),
            spmv64_us=round(ms64 * 1e3, 2),
            spmv64_gbs=round(b64 / (ms64 * 1e-3) / 1e9, 1),
            spmv64_sym_us=None if ms65 is None else round(ms65 * 1e3, 2),
            spmv64_sym_gbs=(None if ms65 is None else
                            round(sym_spmv_bytes(n_nodes, n_valid) / (ms65 * 1e-3) / 1e9, 1)),
            spmv32_us=round(ms32 * 1e3, 2),
            spmv32_gbs=round(spmv_bytes(n_nodes, n_valid, 32) / (ms32 * 1e-3) / 1e9, 1),
            spmv16_us=round(ms16 * 1e3, 2),
            spmv16_gbs=round(spmv_bytes(n_nodes, n_valid, 16) / (ms16 * 1e-3) / 1e9, 1),
            fp64_iter_us=round(1e6 * t_fp64 / iters[-1], 1),
            fp64_iter_gbs=round(pcg_iter_bytes(n_nodes, n_valid, 64, sym) * iters[-1] / t_fp64 / 1e9,
                                1),
            fp64_layout="symmetric" if sym else "full")
    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: ~{args.cpu_seconds}s of iterations, {args.cpu_threads} threads")
        cpu = cpu_baseline(f, p, args.cpu_seconds, args.cpu_threads)
        log(f"[bench] CPU: {cpu['value']:.3e} DOF*it/s ({cpu['kind']}, {cpu['seconds']:.1f}s)")
        cpu["time_to_1e8_fp64_extrapolated_s"] = round(iters[-1] * cpu["seconds"] / cpu["iters"], 1)
        if cpu["one_thread"]:
            o1 = cpu["one_thread"]
            o1["time_to_1e8_fp64_extrapolated_s"] = round(iters[-1] * o1["seconds"] / o1["iters"], 1)
            log(f"[bench] CPU, 1 thread: {o1['value']:.3e} DOF*it/s")
    f.destroy()
    simp = {}
    if not args.no_simp:
        simp = simp_bench(args, solve_opts(MIXED_CHEB_FCG, cheb_degree=args.cheb_degree,
                                           cheb_ratio=args.cheb_ratio,
                                           precond_bits=args.precond_bits, simp=True,
                                           true_residual=True))
    line = dict(
        metric=METRIC, value=value, unit="DOF*CG-iter/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=round(ms_step, 3), higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
        config=dict(workload=cfg["name"], dof=N, elements=p.N_elem, nnz=nnz,
                    stored_blocks=nblk,
                    solver="FP64 Jacobi-PCG from u = 0 to ||F-Ku||/||F|| <= 1e-8",
                    iters_per_step=iters, rel_res=[i["rel_res"] for i in infos],
                    true_rel_res=true64[0], true_rel_res_fp64_eval=true64[1]),
        roofline=roofline, cpu_baseline=cpu, setup_s=dict(pattern=round(t_pattern, 3),
                                                          assembly_bc=round(t_assemble, 3)),
        **extra, **simp)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
