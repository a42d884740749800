"""Benchmark of the nbots FEM solve path on one MI355X.

Workload (default): 2048x1024 cantilever, 4,194,304 T3 elements,
4,200,450 DOF, lognormal(0,1) per-element modulus, clamped left edge,
uniform shear traction on the right edge.  K, F and all solver vectors are
resident in HBM before the timed region.  One step = one FP64
Jacobi-PCG solve (the reference's algorithm) from u = 0 to relative residual
||F - K u|| / ||F|| <= 1e-8.

The JSON line reports DOF * CG-iteration / s over the timed steps (whole
job), the HBM roofline of the dominant kernel (the FP64 SpMV with its p.w
partials, timed with HIP events on the solver's stream), the reference's
own CPU Jacobi-PCG timed on this host over a bounded number of iterations,
and, timed separately on the same GPU, time to 1e-8 of:
  * FP64 Jacobi-PCG (the steps above);
  * FP32 CG sweeps (inner relative tolerance 1e-4) inside FP64 iterative
    refinement, the scheme the workload names;
  * FP64 flexible CG with a degree-40 Chebyshev-Jacobi preconditioner whose
    matrix values, vectors and arithmetic are all FP64;
  * the same with the preconditioner's matrix in fp16 and its vectors in
    FP32 (FP64 row sums): the library's default mixed-precision solver;
  * FP64 flexible CG preconditioned by one FP32 V-cycle of smoothed-
    aggregation multigrid (the hierarchy's setup reported apart).
The preconditioner factor (Jacobi-PCG / FP64-preconditioner FCG) and the
precision factor (FP64-preconditioner FCG / fp16-preconditioner FCG) are
reported apart; their product is the Chebyshev mixed solver's speedup.  A
SIMP topology-optimisation loop on a 1024x512 MBB half-beam (1.05M DOF,
50 iterations) gives ms per SIMP iteration with each solver; the headline
ms_per_simp_iter is the multigrid-preconditioned loop's, setup included.
As context for the roofline fraction, roofline.stream_context gives what
a plain copy kernel (torch's) reaches on the same GPU.

    python bench.py [--steps K] [--warmup W] [--config c3|c2|c4] [--no-cpu]
                    [--simp-iters 50] [--no-simp] [--no-decompose] [--no-stream]
    python bench.py --cpu-full        # the CPU baseline plan at full size
    python bench.py --pmc-probe [--probe-mixed | --probe-amg]   # short runs for rocprofv3
    python bench.py --simp-only [--simp-iters N] [--simp-amg-only]  # the SIMP loop alone
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

METRIC = ("DOF·CG-iter/s + time-to-1e-8 residual (FP32+IR vs FP64); "
          "ms per SIMP iteration")
HBM_PEAK_GBS = 8000.0
# iteration cap of the AMG-preconditioned solves (they take 25-100)
AMG_MAX_ITER = 2000
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


# ---- algorithmic bytes ---------------------------------------------------------

def csr_spmv_bytes(n: int, nnz: int, sv: int = 8) -> int:
    """One SpMV on scalar CSR with 4-byte columns and row pointers: every
    value and column once, the row pointers, x read and y written once."""
    return nnz * (sv + 4) + 4 * (n + 1) + 2 * n * sv


def spmv_bytes(n_nodes: int, n_blocks: int, prec: int) -> int:
    """One 2x2-block sliced-ELL SpMV: per valid block its column (4 B) and 4
    values; per node its row length (1 B) and the x read and y write of 2
    values; per 64-node slice its offset.  prec 64, 32 or 16 (fp16 values
    with FP32 vectors)."""
    sv = {64: 8, 32: 4, 16: 2}[prec]          # matrix value bytes
    sx = 8 if prec == 64 else 4               # vector value bytes (fp16 values: FP32 vectors)
    return n_blocks * (4 * sv + 4) + n_nodes + 4 * ((n_nodes + 63) // 64) + 2 * (2 * n_nodes) * sx


def sym_spmv_bytes(n_nodes: int, n_valid: int) -> int:
    """One FP64 SpMV in the symmetric layout: per node its diagonal block's 3
    values (24 B), two row lengths, the x read and the y write; per upper
    block (v, w > v) its column and 4 values (36 B); per lower block its
    (column, upper position) pair (8 B; the transposed values are those of
    upper blocks streamed by nearby workgroups, served by L2); per 64-node
    slice two offsets.  With every diagonal block stored, upper = lower =
    (valid - nodes) / 2 blocks."""
    up = lo = (n_valid - n_nodes) // 2
    return (n_nodes * 24 + up * 36 + lo * 8 + 2 * n_nodes + 8 * ((n_nodes + 63) // 64)
            + 2 * (2 * n_nodes) * 8)


def cheb_step_bytes(n_nodes: int, n_valid: int, adds_z: bool) -> int:
    """One Chebyshev step on the fp16 copy without diagonal blocks, every
    slice as wide as the widest row (padding not counted): per off-diagonal
    block 8 B of values and a 2-B packed column; per node the diagonal pair
    (4 B), the d gather (8 B), the residual read and write (16 B), the new d
    (8 B) and, on the steps that add to z, its read and write (16 B)."""
    off = n_valid - n_nodes
    per_node = 4 + 8 + 16 + 8 + (16 if adds_z else 0)
    return off * 10 + n_nodes * per_node


def pcg_iter_bytes(n_nodes, n_blocks, prec, sym=False):
    """One CG iteration as implemented: SpMV kernel (above) + update kernel
    (FP64: read g,w,diag, write g = 4 vectors; FP32 scaled: read y,p,r,w,
    write y,r = 6) + direction kernel (FP64: read p,g,diag and write the
    next p = 4 vectors, and on every fourth iteration also read the three
    previous p and read and write x = 9; FP32: read r,p, write p = 3)."""
    sv = 8 if prec == 64 else 4
    nv = 9.25 if prec == 64 else 9
    spmv = sym_spmv_bytes(n_nodes, n_blocks) if sym else spmv_bytes(n_nodes, n_blocks, prec)
    return spmv + nv * (2 * n_nodes) * sv


# ---- host description ---------------------------------------------------------

def host_info():
    """nproc, the CPUs this process may run on, OMP_NUM_THREADS and the CPU
    model.  A shared GPU host allots each GPU's job a share of its cores
    (the affinity set and OMP_NUM_THREADS say how many)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    return dict(nproc=os.cpu_count(), affinity=aff,
                omp_num_threads=int(omp) if omp and omp.isdigit() else None, cpu_model=model)


def default_cpu_threads():
    h = host_info()
    n = h["affinity"]
    if h["omp_num_threads"]:
        n = min(n, h["omp_num_threads"])
    return max(1, n)


# ---- CPU baseline -----------------------------------------------------------------

def _ref_pcg_runner(p, rp, col, val, F):
    """(run(n_iter, threads) -> seconds, kind, release): the reference's
    nb_sparse_solve_CG_precond_Jacobi compiled from its own sources into
    oracle/_ref, or this repository's C restatement of it (1 thread)."""
    import ctypes as C
    import oracle
    it, tr = C.c_uint32(0), C.c_double(0)
    try:
        R = oracle.ref()
    except Exception:
        R = None
    if R is not None:
        h = R.ref_sparse_import(p.N, rp, col, val)

        def run(n, nt, tol=0.0):
            x = np.zeros(p.N)
            t0 = time.perf_counter()
            R.ref_pcg_handle(h, F, x, n, tol, nt, C.byref(it), C.byref(tr))
            return time.perf_counter() - t0, it.value, x
        return run, "reference", lambda: R.ref_sparse_free(h)

    def run(n, nt, tol=0.0):
        x = np.zeros(p.N)
        t0 = time.perf_counter()
        oracle.lib().orc_pcg_jacobi(p.N, rp, col, val, F, x, n, tol, C.byref(it), C.byref(tr))
        return time.perf_counter() - t0, it.value, x
    return run, "port", lambda: None


def cpu_baseline(f, p, seconds, threads):
    """The reference's Jacobi-PCG on the same K and F, cut after about
    `seconds` of work (10 calibration iterations decide the count), on
    `threads` OpenMP threads and again on 1 (the reference FEM driver's own
    setting, static_elasticity2D.c:88-90) for a quarter of the time."""
    rp, col, val, F = f.system()
    run, kind, release = _ref_pcg_runner(p, rp, col, val, F)
    if kind == "port":
        threads = 1

    def sample(secs, nt):
        t_cal, _, _ = run(10, nt)
        n = int(max(10, min(20000, secs / max(t_cal / 10, 1e-6))))
        dt, n_it, _ = run(n, nt)
        return n_it, dt

    n_it, dt = sample(seconds, threads)
    one = None
    if kind == "reference" and threads > 1:
        n1, dt1 = sample(seconds / 4, 1)
        one = dict(value=p.N * n1 / dt1, iters=n1, seconds=dt1)
    release()
    return dict(value=p.N * n_it / dt, unit="DOF*CG-iter/s", cores=threads, kind=kind,
                sample=f"{n_it} Jacobi-PCG iterations on the same {p.N}-DOF system "
                       f"(reference build gcc -O2, {threads} OpenMP threads), {dt:.2f} s; "
                       f"time to 1e-8 extrapolated from the FP64 GPU iteration count",
                seconds=dt, iters=n_it, one_thread=one, host=host_info())


def cpu_full(args):
    """The CPU baseline plan at full size, on this host's cores (no GPU):
    C1 64x32 jittered cantilever, the reference driver's solve (absolute
    tolerance 1e-8, 1 thread); C2 1024x512: the reference's pattern,
    assembly and BCs, then its Jacobi-PCG to relative 1e-8 on `threads` and
    on 1 thread; C5 1024x512 MBB from rho = 0.5: one SIMP iteration (the
    reference's assembly with E(xPhys) and its Jacobi-PCG to relative 1e-8,
    then this repository's C restatement of the element energies,
    sensitivities, filter and OC update)."""
    import oracle
    from nbots_amd import synth
    threads = args.cpu_threads or default_cpu_threads()
    out = dict(host=host_info(), threads=threads, build="gcc -O2 -std=c99 -fopenmp "
               "-ffp-contract=off (reference sources, oracle/Makefile)")

    def timed(fn, *a, **k):
        t0 = time.perf_counter()
        r = fn(*a, **k)
        return r, time.perf_counter() - t0

    # C1
    p = synth.cantilever_point(64, 32, jitter=0.1)
    res, t = timed(oracle.ref_fem_compute, p)
    out["c1"] = dict(workload="64x32 jittered cantilever, 4,290 DOF, the reference driver "
                              "(pattern, assembly, BCs, PCG to absolute 1e-8, strain), 1 thread",
                     seconds=round(t, 4), status=int(res[0]))
    log(f"[cpu-full] C1 driver {t:.3f}s")
    # C2
    p = synth.cantilever_traction(1024, 512, seed=args.seed)
    rp, col, val, F, st, t_pat, t_asm = oracle.ref_assemble_timed(p)
    assert st == 0
    tol = 1e-8 * float(np.linalg.norm(F))
    run, kind, release = _ref_pcg_runner(p, rp, col, val, F)
    c2 = dict(workload=f"1024x512 lognormal cantilever, {p.N} DOF", kind=kind,
              pattern_s=round(t_pat, 3), pattern_assembly_bc_s=round(t_asm, 3),
              assembly_bc_s=round(t_asm - t_pat, 3),
              note="pattern = mesh graph + nb_sparse_create + CSR export; assembly with "
                   "per-element moduli through the reference's element kernels")
    for nt in sorted({threads, 1}, reverse=True):
        dt, it, x = run(p.N, nt, tol)
        rel = float(np.linalg.norm(F - oracle.ref_spmv(rp, col, val, x, threads=nt)) /
                    np.linalg.norm(F))
        c2[f"solve_{nt}t"] = dict(seconds=round(dt, 3), niter=it, rel_res=rel,
                                  dof_iter_per_s=p.N * it / dt)
        log(f"[cpu-full] C2 PCG {nt} threads: {it} it in {dt:.1f}s (rel res {rel:.2e})")
    release()
    out["c2"] = c2
    # C5, one SIMP iteration
    p = synth.mbb(1024, 512, rho=0.5)
    cx, cy = synth.elem_centroids(p)
    H, t_filter = timed(oracle.filter_build, cx, cy, 1.5)
    x = np.full(p.N_elem, 0.5)
    xp = oracle.filter_apply(H, x)
    q = synth.mbb(1024, 512, rho=xp)
    rp, col, val, F, st, _, t_asm = oracle.ref_assemble_timed(q)
    tol = 1e-8 * float(np.linalg.norm(F))
    run, kind, release = _ref_pcg_runner(q, rp, col, val, F)
    t_solve, it, u = run(q.N * 20, threads, tol)
    release()
    t0 = time.perf_counter()
    ce = oracle.elem_energy(q, u)
    C_, dc = oracle.simp_sens(xp, ce)
    dcf = oracle.filter_sens(H, dc)
    dvf = oracle.filter_sens(H, synth.elem_areas(q))
    xnew, _, ch = oracle.oc_update(H, synth.elem_areas(q), x, dcf, dvf, 0.5, 0.2)
    t_sens = time.perf_counter() - t0
    out["c5"] = dict(workload=f"1024x512 MBB half-beam from rho = 0.5, {q.N} DOF, first SIMP "
                              "iteration", kind=kind,
                     assembly_bc_s=round(t_asm, 3), solve_s=round(t_solve, 3), solve_niter=it,
                     solve_threads=threads,
                     sens_filter_oc_s=round(t_sens, 3), sens_kind="port (1 thread)",
                     filter_build_s=round(t_filter, 3),
                     ms_per_simp_iter=round(1e3 * (t_asm + t_solve + t_sens), 1),
                     compliance=float(C_))
    log(f"[cpu-full] C5 SIMP iteration: assembly {t_asm:.2f}s, solve {t_solve:.1f}s "
        f"({it} it), sens/filter/OC {t_sens:.2f}s")
    print(json.dumps(out), flush=True)


# ---- counter probe ------------------------------------------------------------------

def pmc_probe(args):
    """A short FP64 Jacobi-PCG run (or, with --probe-mixed, outer iterations
    of the mixed solver) on the cantilever, then a plain elementwise kernel
    of known size, for hardware-counter runs:

      rocprofv3 --pmc FETCH_SIZE --kernel-include-regex 'k_pcg_A|elementwise' \\
          -d <dir> -- python bench.py --pmc-probe

    The elementwise b = 2a reads and writes exactly --copy-mib MiB, so its
    counter reading calibrates the counter for the solver's kernels."""
    import torch
    from nbots_amd import synth
    from nbots_amd.fem import Fem, SOLVE_FP64, SOLVE_MIXED
    cfg = CONFIGS[args.config]
    p = synth.cantilever_traction(cfg["nx"], cfg["ny"], seed=args.seed)
    f = Fem.from_problem(p)
    N, nnz, nblk = f.sizes()
    n_nodes, n_valid = N // 2, nnz // 4
    print(f"[probe] {cfg['nx']}x{cfg['ny']} cantilever: {N} DOF, {n_valid} blocks, "
          f"SpMV algorithmic bytes {spmv_bytes(n_nodes, n_valid, 64)} (FP64, full layout), "
          f"{sym_spmv_bytes(n_nodes, n_valid)} (symmetric layout)", flush=True)
    if args.probe_amg:
        from nbots_amd.fem import MIXED_AMG_FCG
        o = _opts(MIXED_AMG_FCG, true_residual=True)
        for rep in range(3):
            t0 = time.perf_counter()
            info = f.solve(SOLVE_MIXED, tol=1e-8, max_iter=AMG_MAX_ITER, opts=o)
            print(f"[probe] AMG-preconditioned FCG {rep} ({'with' if rep == 0 else 'without'} "
                  f"setup): {info['iters']} iterations, {time.perf_counter() - t0:.3f} s, "
                  f"layout {info['layout']}", flush=True)
    elif args.probe_mixed:
        print(f"[probe] Chebyshev step algorithmic bytes {cheb_step_bytes(n_nodes, n_valid, True)} "
              f"(adding to z), {cheb_step_bytes(n_nodes, n_valid, False)} (not)", flush=True)
        info = f.solve(SOLVE_MIXED, tol=0.0, max_iter=args.probe_iters)
        print(f"[probe] mixed: {info['iters']} outer iterations, layout {info['layout']}", flush=True)
    else:
        info = f.solve(SOLVE_FP64, tol=0.0, max_iter=args.probe_iters)
        print(f"[probe] FP64 PCG: {info['iters']} iterations, layout {info['layout']}", flush=True)
    f.destroy()
    # b = 2 a: an elementwise kernel that reads and writes every byte once
    # (a same-type copy_ would go to the runtime's blit path instead)
    n = args.copy_mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    print(f"[probe] elementwise b = 2a: 3 launches of {n * 4} bytes read + {n * 4} bytes written",
          flush=True)


def stream_rates(mib=1024, reps=20):
    """What plain streaming kernels reach on this GPU, as context for the
    SpMV's fraction of the 8 TB/s peak: torch's b.copy_(a) over `mib` MiB
    (read + write), the median of `reps` launches timed with events on
    torch's stream (torch's reductions stream slower still).  Runs in
    a process of its own (`--stream-only`): torch brings its own HIP runtime,
    which cannot start in a process where the library's has."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty(n, dtype=torch.float32, device="cuda")

    def med(fn, nbytes):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return round(nbytes / float(np.median(ts)) / 1e9, 1)

    out = dict(copy_gbs=med(lambda: b.copy_(a), 8 * n),
               note=f"torch b.copy_(a) of {mib} MiB (read and written), median of {reps} "
                    "launches")
    return out


def stream_rates_child():
    """stream_rates() in a child process; None when it fails."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--stream-only"],
                       capture_output=True, text=True, timeout=300)
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        log(f"[bench] streaming context failed (exit {r.returncode}): {r.stderr[-400:]}")
        return None
    return json.loads(lines[-1])


# ---- GPU solvers --------------------------------------------------------------------

REPLAY = False   # --replay: the solvers' repeated steps as captured HIP graphs


def _opts(*a, **kw):
    """nbx_solve_opts_t for every solve of this run (see nbots_amd.fem.solve_opts)."""
    from nbots_amd.fem import solve_opts
    return solve_opts(*a, replay=REPLAY, **kw)


def time_solver(f, mode, opts, reps, label, max_iter=200000):
    """One untimed solve, then `reps` timed ones; returns (median seconds,
    last info)."""
    f.solve(mode, tol=1e-8, max_iter=max_iter, opts=opts)
    ts, info = [], None
    for s in range(reps):
        t0 = time.perf_counter()
        info = f.solve(mode, tol=1e-8, max_iter=max_iter, opts=opts)
        ts.append(time.perf_counter() - t0)
        log(f"[bench] {label} {s}: {info['iters']} it, {info['outer']} inner, {ts[-1]:.3f}s, "
            f"rel_res {info['rel_res']:.2e}, fallback {info['fallback']}, status {info['status']}")
    return float(np.median(ts)), info


def amg_solve(f, t_fp64, u64, Fl):
    """FP64 flexible CG preconditioned by one FP32 smoothed-aggregation
    V-cycle.  The first solve builds the hierarchy (host symbolic phase +
    device numeric phase); the timed solves reuse it, as solves of one matrix
    with several loads do.  A SIMP iteration pays the numeric phase again
    (the moduli change), which the SIMP legs include."""
    from nbots_amd.fem import SOLVE_MIXED, MIXED_AMG_FCG
    o = _opts(MIXED_AMG_FCG, true_residual=True)
    t0 = time.perf_counter()
    first = f.solve(SOLVE_MIXED, tol=1e-8, max_iter=AMG_MAX_ITER, opts=o)
    t_first = time.perf_counter() - t0
    log(f"[bench] AMG first solve (with setup): {first['iters']} it, {t_first:.3f}s, "
        f"status {first['status']}")
    t, info = time_solver(f, SOLVE_MIXED, o, 3, "AMG-preconditioned FCG", max_iter=AMG_MAX_ITER)
    u = f.displacement()
    return dict(
        time_to_1e8_amg_s=round(t, 4),
        amg=dict(seconds=round(t, 4), speedup_vs_fp64=round(t_fp64 / t, 3),
                 iters=info["iters"], status=info["status"], rel_res=info["rel_res"],
                 true_rel_res=f.true_residual(True),
                 first_solve_s=round(t_first, 4), setup_s=round(t_first - t, 4),
                 vs_fp64_disp_rel_l2=float(np.linalg.norm(u - u64) / np.linalg.norm(u64)),
                 vs_fp64_compliance_rel=abs(float(Fl @ u) - float(Fl @ u64)) / abs(float(Fl @ u64)),
                 solver=("FP64 flexible CG; preconditioner one FP32 V-cycle of smoothed-"
                         "aggregation multigrid (rigid-body modes; one weighted Jacobi step on "
                         "the finest level's fp16 scaled copy, Chebyshev-Jacobi smoothing of "
                         "degree 2, 4 and 5 below; dense coarsest level); setup = first solve - solve")))


def decompose(f, args, t_fp64, iters64, u64, Fl):
    """The mixed-precision speedup split into the preconditioner's gain and
    the precision's, plus the FP32-CG + FP64-refinement scheme."""
    from nbots_amd.fem import SOLVE_MIXED, MIXED_CHEB_FCG, MIXED_CG_RESTART
    out = {}
    o16 = _opts(MIXED_CHEB_FCG, cheb_degree=args.cheb_degree, cheb_ratio=args.cheb_ratio,
                     precond_bits=args.precond_bits, true_residual=True)
    t16, i16 = time_solver(f, SOLVE_MIXED, o16, 2, "mixed (fp16 preconditioner)")
    u16 = f.displacement()
    true16 = (f.true_residual(True), f.true_residual(False))
    out.update(
        time_to_1e8_fp64_s=round(t_fp64, 4),
        time_to_1e8_mixed_s=round(t16, 4),
        mixed_speedup_vs_fp64=round(t_fp64 / t16, 3),
        mixed_solver=("FP64 flexible CG, Chebyshev-Jacobi preconditioner (degree "
                      f"{o16.cheb_degree} on [lmax/{o16.cheb_ratio:g}, lmax]) on "
                      f"{'fp16' if o16.precond_bits == 16 else 'FP32'} matrix values, FP32 vectors, "
                      "FP64 row sums"),
        mixed_precond_matrix=(("fp16 copy without diagonal blocks"
                               + (", 16-bit column deltas" if i16["layout"] & 4 else ""))
                              if i16["layout"] & 2 else "full scaled copy"),
        mixed_iters=i16["iters"], mixed_inner=i16["outer"], mixed_rel_res=i16["rel_res"],
        mixed_true_rel_res=true16[0], mixed_true_rel_res_fp64_eval=true16[1],
        mixed_vs_fp64_disp_rel_l2=float(np.linalg.norm(u16 - u64) / np.linalg.norm(u64)),
        mixed_vs_fp64_compliance_rel=abs(float(Fl @ u16) - float(Fl @ u64)) / abs(float(Fl @ u64)))
    out.update(amg_solve(f, t_fp64, u64, Fl))
    if args.no_decompose:
        return out
    o64 = _opts(MIXED_CHEB_FCG, cheb_degree=args.cheb_degree, cheb_ratio=args.cheb_ratio,
                     precond_bits=64, true_residual=True)
    t64, i64 = time_solver(f, SOLVE_MIXED, o64, 2, "FCG, FP64 preconditioner")
    true64p = f.true_residual(True)
    oir = _opts(MIXED_CG_RESTART, inner_rtol=1e-4, true_residual=True)
    tir, iir = time_solver(f, SOLVE_MIXED, oir, 1, "FP32 CG + FP64 refinement",
                           max_iter=args.mixed_max_iter)
    trueir = f.true_residual(True)
    out.update(
        decomposition=dict(
            jacobi_pcg_fp64=dict(seconds=round(t_fp64, 4), iters=iters64),
            cheb_fcg_fp64_precond=dict(
                seconds=round(t64, 4), iters=i64["iters"], precond_spmvs=i64["outer"],
                true_rel_res=true64p,
                solver="FP64 flexible CG, Chebyshev-Jacobi preconditioner on the FP64 matrix, "
                       "FP64 vectors and arithmetic"),
            cheb_fcg_fp16_precond=dict(seconds=round(t16, 4), iters=i16["iters"],
                                       precond_spmvs=i16["outer"], true_rel_res=true16[0]),
            preconditioner_factor=round(t_fp64 / t64, 3),
            precision_factor=round(t64 / t16, 3),
            note="mixed_speedup_vs_fp64 = preconditioner_factor x precision_factor"),
        time_to_1e8_fp32ir_s=round(tir, 4),
        fp32ir=dict(seconds=round(tir, 4), speedup_vs_fp64=round(t_fp64 / tir, 3),
                    inner_fp32_iters=iir["iters"], refinement_sweeps=iir["outer"],
                    fp64_pcg_iters=iir["fp64_iters"], fallback_to_fp64=bool(iir["fallback"]),
                    status=iir["status"], rel_res=iir["rel_res"], true_rel_res=trueir,
                    solver="FP32 CG sweeps to inner relative residual 1e-4 on the FP32 "
                           "Jacobi-scaled matrix, FP64 residual and update; a sweep whose "
                           "true FP64 residual does not fall hands over to FP64 PCG"))
    return out


def simp_bench(args, opts16):
    """1024x512 MBB half-beam (1,051,650 DOF), volume fraction 0.5, penalty 3,
    density filter radius 1.5, OC move 0.2, from rho = 0.5: the SIMP loop
    with each solver (all to relative residual 1e-8, warm-started from the
    previous displacement): the flexible CG preconditioned by one FP32
    smoothed-aggregation V-cycle, the flexible CG with the fp16 Chebyshev
    preconditioner, the same with an FP64 preconditioner, FP64 Jacobi-PCG,
    and, for the first --simp-fp32ir-iters iterations, FP32 CG in FP64
    refinement.  With --simp-amg-only, the multigrid loop alone."""
    from nbots_amd import synth
    from nbots_amd.fem import (Fem, Simp, SOLVE_FP64, SOLVE_MIXED, MIXED_CHEB_FCG,
                               MIXED_CG_RESTART, MIXED_AMG_FCG)
    nx, ny = 1024, 512
    p = synth.mbb(nx, ny, rho=0.5)
    out = dict(simp_workload=f"{nx}x{ny} MBB half-beam, {p.N} DOF, {p.N_elem} T3, volfrac 0.5, "
                             "penal 3, filter radius 1.5, OC move 0.2, rho0 = 0.5",
               simp_iterations=args.simp_iters)
    variants = [("amg", SOLVE_MIXED, _opts(MIXED_AMG_FCG, simp=True, true_residual=True),
                 args.simp_iters),
                ("mixed", SOLVE_MIXED, opts16, args.simp_iters),
                ("fp64", SOLVE_FP64, _opts(true_residual=True, simp=True), args.simp_iters)]
    if args.simp_amg_only:
        variants = variants[:1]
    elif not args.no_decompose:
        variants.insert(1, ("fcg_fp64_precond", SOLVE_MIXED,
                            _opts(MIXED_CHEB_FCG, precond_bits=64, simp=True,
                                       true_residual=True), args.simp_iters))
        if args.simp_fp32ir_iters > 0:
            variants.append(("fp32ir", SOLVE_MIXED,
                             _opts(MIXED_CG_RESTART, inner_rtol=1e-4, simp=True,
                                        true_residual=True), args.simp_fp32ir_iters))
    runs = {}
    for name, mode, o, n_it in variants:
        f = Fem.from_problem(p)
        t0 = time.perf_counter()
        s = Simp(f, rmin=1.5, volfrac=0.5, penal=3.0, E0=1.0, Emin=1e-9, move=0.2)
        t_setup = time.perf_counter() - t0
        hist, ch, its, solve_ms, it_s, fb = [], [], [], [], [], 0
        for it in range(n_it):
            t0 = time.perf_counter()
            C, change, info = s.iterate(mode, tol=1e-8, opts=o,
                                        max_iter=AMG_MAX_ITER if name == "amg" else 0)
            it_s.append(time.perf_counter() - t0)
            hist.append(C)
            ch.append(change)
            its.append(info["iters"])
            solve_ms.append(info["ms"])
            fb += int(info["fallback"] != 0 and name == "fp32ir")
            if it % 10 == 0 or it == n_it - 1:
                log(f"[bench] SIMP {name} it {it}: C {C:.6e} change {change:.4f} "
                    f"solve {info['iters']} it {info['ms']:.1f} ms rel_res {info['rel_res']:.1e}")
        runs[name] = dict(hist=np.array(hist), it_s=np.array(it_s), solve_ms=float(np.mean(solve_ms)),
                          iters=its, setup_s=t_setup, change=ch, fallbacks=fb)
        s.destroy()
        f.destroy()
    if args.simp_amg_only:
        a = runs["amg"]
        out.update(ms_per_simp_iter=round(1e3 * float(a["it_s"].mean()), 2),
                   simp_solve_ms_mean=round(a["solve_ms"], 2),
                   simp_solver_iters_mean=round(float(np.mean(a["iters"])), 1),
                   simp_amg_first_iter_ms=round(1e3 * float(a["it_s"][0]), 2),
                   simp_compliance_first_last=[float(a["hist"][0]), float(a["hist"][-1])])
        return out
    a, m, d = runs["amg"], runs["mixed"], runs["fp64"]
    rel = np.abs(m["hist"] - d["hist"]) / np.abs(d["hist"])
    rel_a = np.abs(a["hist"] - d["hist"]) / np.abs(d["hist"])
    ms = {k: 1e3 * float(v["it_s"].mean()) for k, v in runs.items()}
    out.update(
        ms_per_simp_iter=round(ms["amg"], 2),
        simp_solver="FP64 flexible CG with one FP32 smoothed-aggregation V-cycle as preconditioner",
        ms_per_simp_iter_cheb=round(ms["mixed"], 2),
        ms_per_simp_iter_fp64=round(ms["fp64"], 2),
        simp_solve_ms_mean=round(a["solve_ms"], 2),
        simp_solve_ms_mean_cheb=round(m["solve_ms"], 2),
        simp_solve_ms_mean_fp64=round(d["solve_ms"], 2),
        simp_filter_setup_s=round(m["setup_s"], 3),
        simp_solver_iters_mean=round(float(np.mean(a["iters"])), 1),
        simp_solver_iters_mean_cheb=round(float(np.mean(m["iters"])), 1),
        simp_solver_iters_mean_fp64=round(float(np.mean(d["iters"])), 1),
        simp_amg_first_iter_ms=round(1e3 * float(a["it_s"][0]), 2),
        simp_compliance_first_last=[float(d["hist"][0]), float(d["hist"][-1])],
        simp_compliance_hist_max_rel_diff=float(rel_a.max()),
        simp_compliance_hist_max_rel_diff_cheb=float(rel.max()),
        simp_change_last=float(a["change"][-1]))
    if "fcg_fp64_precond" in runs:
        q = runs["fcg_fp64_precond"]
        out["simp_decomposition"] = dict(
            ms_per_iter=dict(jacobi_pcg_fp64=round(ms["fp64"], 2),
                             cheb_fcg_fp64_precond=round(ms["fcg_fp64_precond"], 2),
                             cheb_fcg_fp16_precond=round(ms["mixed"], 2)),
            preconditioner_factor=round(ms["fp64"] / ms["fcg_fp64_precond"], 3),
            precision_factor=round(ms["fcg_fp64_precond"] / ms["mixed"], 3),
            fcg_fp64_precond_hist_max_rel_diff=float(
                (np.abs(q["hist"] - d["hist"]) / np.abs(d["hist"])).max()))
    if "fp32ir" in runs:
        r = runs["fp32ir"]
        n = len(r["it_s"])
        first = {k: round(1e3 * float(v["it_s"][:n].mean()), 2) for k, v in runs.items()}
        out["simp_fp32ir"] = dict(
            iterations=n, ms_per_iter_first_n=first,
            solver_iters_mean=round(float(np.mean(r["iters"])), 1),
            solves_handed_to_fp64=r["fallbacks"],
            hist_max_rel_diff_vs_fp64=float(
                (np.abs(r["hist"] - d["hist"][:n]) / np.abs(d["hist"][:n])).max()),
            note=f"ms per iteration over the first {n} iterations for every solver")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="default: the cores this job may use (affinity, OMP_NUM_THREADS)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="the CPU baseline plan at full size instead (no GPU)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="a short solve and a calibration kernel for rocprofv3 --pmc")
    ap.add_argument("--probe-mixed", action="store_true")
    ap.add_argument("--probe-amg", action="store_true",
                    help="with --pmc-probe: three AMG-preconditioned solves (the first builds "
                         "the hierarchy)")
    ap.add_argument("--probe-iters", type=int, default=50)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the plain streaming kernels timed as roofline context")
    ap.add_argument("--stream-only", action="store_true",
                    help="only those streaming kernels, as one JSON line")
    ap.add_argument("--no-decompose", action="store_true",
                    help="skip the FP64-preconditioner and FP32-CG + FP64-refinement runs")
    ap.add_argument("--mixed-max-iter", type=int, default=200000)
    ap.add_argument("--cheb-degree", type=int, default=None, help="library default 40")
    ap.add_argument("--cheb-ratio", type=float, default=None, help="library default 3e4")
    ap.add_argument("--precond-bits", type=int, default=16, choices=[16, 32])
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--simp-iters", type=int, default=50)
    ap.add_argument("--simp-fp32ir-iters", type=int, default=10)
    ap.add_argument("--no-simp", action="store_true")
    ap.add_argument("--replay", action="store_true",
                    help="replay the solvers' repeated steps as captured HIP graphs instead of "
                         "launching kernel by kernel (same results)")
    ap.add_argument("--simp-amg-only", action="store_true",
                    help="the SIMP loop with multigrid-preconditioned solves only")
    ap.add_argument("--simp-only", action="store_true",
                    help="only the SIMP loop, its results as one JSON line (for profiling)")
    args = ap.parse_args()
    global REPLAY
    REPLAY = args.replay
    if args.gpus != 1:
        print(f"bench.py measures one MI355X only (asked for --gpus {args.gpus}); "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.cpu_full:
        cpu_full(args)
        return
    if args.pmc_probe:
        pmc_probe(args)
        return
    if args.stream_only:
        print(json.dumps(stream_rates()), flush=True)
        return
    if args.simp_only:
        from nbots_amd.fem import MIXED_CHEB_FCG
        print(json.dumps(simp_bench(args, _opts(
            MIXED_CHEB_FCG, cheb_degree=args.cheb_degree, cheb_ratio=args.cheb_ratio,
            precond_bits=args.precond_bits, simp=True, true_residual=True))), flush=True)
        return

    from nbots_amd import synth
    from nbots_amd.fem import Fem, SOLVE_FP64, MIXED_CHEB_FCG

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
    # The solvers stop on the true residual ||F - K u|| / ||F|| <= 1e-8
    # (restarting from the iterate when their recursively updated residual
    # has drifted from it).
    o64 = _opts(true_residual=True)
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
    t_fp64 = float(np.median(times))

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
                    bytes_per_launch=bk, us_per_launch=round(spmv_ms * 1e3, 2),
                    byte_model=("symmetric 2x2-block sliced ELL: 24 B per node for the "
                                "diagonal block, 36 B per upper block (column + 4 FP64 values), "
                                "8 B per lower block (column + position of its upper partner, "
                                "values read transposed from L2), x and y 16 B per node each"
                                if sym else
                                "2x2-block sliced ELL: 36 B per block, x and y 16 B per node"),
                    csr_bytes_per_launch=csr_spmv_bytes(N, nnz),
                    csr_note="the same product on scalar CSR with 4-byte indices (the survey's "
                             "byte model); this layout moves fewer bytes for the same result")
    st = None if args.no_stream else stream_rates_child()
    if st:
        st["spmv_frac_of_copy"] = round(achieved / st["copy_gbs"], 3)
        roofline["stream_context"] = st
        log(f"[bench] streaming context: torch copy {st['copy_gbs']} GB/s")

    extra = {}
    if not args.no_mixed:
        extra = decompose(f, args, t_fp64, iters[-1], u64, f.load_vector())
        ms32 = f.spmv_ms(32, 20)
        ms16 = f.spmv_ms(16, 20)
        ms64 = f.spmv_ms(64, 20)
        ms65 = f.spmv_ms(65, 20) if sym else None
        extra.update(
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
        threads = args.cpu_threads or default_cpu_threads()
        log(f"[bench] CPU baseline: ~{args.cpu_seconds}s of iterations, {threads} threads")
        cpu = cpu_baseline(f, p, args.cpu_seconds, threads)
        log(f"[bench] CPU: {cpu['value']:.3e} DOF*it/s ({cpu['kind']}, {cpu['seconds']:.1f}s)")
        cpu["time_to_1e8_fp64_extrapolated_s"] = round(iters[-1] * cpu["seconds"] / cpu["iters"], 1)
        if cpu["one_thread"]:
            o1 = cpu["one_thread"]
            o1["time_to_1e8_fp64_extrapolated_s"] = round(iters[-1] * o1["seconds"] / o1["iters"], 1)
            log(f"[bench] CPU, 1 thread: {o1['value']:.3e} DOF*it/s")
    f.destroy()
    simp = {}
    if not args.no_simp:
        simp = simp_bench(args, _opts(MIXED_CHEB_FCG, cheb_degree=args.cheb_degree,
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
