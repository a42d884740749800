"""Counter probe: a short FP64 Jacobi-PCG run on the cantilever system and a
plain device copy of known size, for hardware-counter runs of rocprofv3.

    rocprofv3 --pmc FETCH_SIZE --kernel-include-regex 'k_pcg_A|elementwise' \\
        -d <dir> -- python -m nbots_amd.probe --nx 2048 --ny 1024 --iters 50

The SpMV kernel of every CG iteration (k_pcg_A) streams the 2x2-block
matrix once and gathers x; the copy reads and writes exactly --copy-mib MiB,
so the ratio of the copy's counter reading to its size calibrates the
counter for the SpMV.  The probe prints the algorithmic bytes of one SpMV
launch for comparison.  No timing is taken here.

--mixed runs --iters outer iterations of the default mixed-precision solve
instead, for the counters of its Chebyshev steps (k_cheb_step64).
"""
from __future__ import annotations

import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def spmv_bytes(n_nodes: int, n_blocks: int, prec: int) -> int:
    """Algorithmic HBM bytes of one 2x2-block sliced-ELL SpMV: per valid
    block its column (4 B) and 4 values; per node its row length (1 B) and
    the x read and y write of 2 values; per 64-node slice its offset.
    prec 64, 32 or 16 (fp16 values with FP32 vectors)."""
    sv = {64: 8, 32: 4, 16: 2}[prec]          # matrix value bytes
    sx = 8 if prec == 64 else 4               # vector value bytes (fp16 values: FP32 vectors)
    return n_blocks * (4 * sv + 4) + n_nodes + 4 * ((n_nodes + 63) // 64) + 2 * (2 * n_nodes) * sx


def sym_spmv_bytes(n_nodes: int, n_valid: int) -> int:
    """Algorithmic HBM bytes of one FP64 SpMV in the symmetric layout: per
    node its diagonal block's 3 values (24 B), two row lengths, the x read
    and the y write; per upper block (v, w > v) its column and 4 values
    (36 B); per lower block its (column, upper position) pair (8 B; the
    transposed values are those of upper blocks streamed by nearby
    workgroups, served by L2); per 64-node slice two offsets.  With every
    diagonal block stored, upper = lower = (valid - nodes) / 2 blocks."""
    up = lo = (n_valid - n_nodes) // 2
    return (n_nodes * 24 + up * 36 + lo * 8 + 2 * n_nodes + 8 * ((n_nodes + 63) // 64)
            + 2 * (2 * n_nodes) * 8)


def cheb_step_bytes(n_nodes: int, n_valid: int, adds_z: bool) -> int:
    """Algorithmic bytes of one Chebyshev step on the fp16 copy without
    diagonal blocks: per off-diagonal block 8 B of values and a 2-B packed
    column; per node the diagonal pair (4 B), the row length (1 B), the d
    gather (8 B), the residual read and write (16 B), the new d (8 B) and,
    on the steps that add to z, its read and write (16 B); per 64-node slice
    two offsets."""
    off = n_valid - n_nodes
    per_node = 4 + 1 + 8 + 16 + 8 + (16 if adds_z else 0)
    return off * 10 + n_nodes * per_node + 8 * ((n_nodes + 63) // 64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--mixed", action="store_true")
    args = ap.parse_args()

    import torch
    from nbots_amd import synth
    from nbots_amd.fem import Fem, SOLVE_FP64, SOLVE_MIXED

    p = synth.cantilever_traction(args.nx, args.ny, seed=args.seed)
    f = Fem.from_problem(p)
    N, nnz, nblk = f.sizes()
    n_nodes, n_valid = N // 2, nnz // 4
    print(f"[probe] {args.nx}x{args.ny} cantilever: {N} DOF, {n_valid} blocks, "
          f"SpMV algorithmic bytes {spmv_bytes(n_nodes, n_valid, 64)} (FP64, full layout), "
          f"{sym_spmv_bytes(n_nodes, n_valid)} (symmetric layout)", flush=True)
    if args.mixed:
        print(f"[probe] Chebyshev step algorithmic bytes {cheb_step_bytes(n_nodes, n_valid, True)} "
              f"(adding to z), {cheb_step_bytes(n_nodes, n_valid, False)} (not)", flush=True)
        info = f.solve(SOLVE_MIXED, tol=0.0, max_iter=args.iters)
        print(f"[probe] mixed: {info['iters']} outer iterations, layout {info['layout']}", flush=True)
    else:
        info = f.solve(SOLVE_FP64, tol=0.0, max_iter=args.iters)
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


if __name__ == "__main__":
    main()
