"""Kernel profiling driver for rocprofv3.

Runs, on the hex8 elasticity workload (default 96^3 elements), a fixed
sequence of launches so that kernel traces and PMC counters can be
attributed per kernel:
  reps x plain device copy of 1 GiB       (bhip_bench_stream: k_stream_copy)
  reps x FP32 SELL-64 3x3-block SpMV      (k_sell_spmv<3, 1>)
  reps x FP64 CSR SpMV                    (k_csr_spmv<8,double,double,1>)
  one FP32-inner + FP64-refined solve     (unless --no-solve)
and prints the HIP-event timings of each group.

usage: python -m betanoox_amd.profile_kernels [--n 96] [--reps 10] [--no-solve]
"""
import argparse
import json
import sys

from . import capi, meshgen, setup_problem


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--no-solve", action="store_true")
    a = ap.parse_args(argv)
    prob = meshgen.elastic_cube(a.n)
    out = {"n": a.n, "n_dof": prob.n_dof}
    with capi.Context(0) as ctx:
        setup_problem(ctx, prob)
        ctx.assemble()
        ctx.set_load(prob.load)
        ctx.build_rhs()
        ms, gbps = ctx.bench_stream(1 << 30, a.reps)
        out["stream_copy"] = {"bytes_read_plus_written": 2 * (1 << 30), "ms": ms, "GBps": gbps}
        ms, by = ctx.bench_spmv(fp32=True, reps=a.reps)
        out["sell3_fp32"] = {"bytes": by, "ms": ms, "GBps": by / ms / 1e6}
        ms, by = ctx.bench_spmv(fp32=False, reps=a.reps)
        out["csr_fp64"] = {"bytes": by, "ms": ms, "GBps": by / ms / 1e6}
        if not a.no_solve:
            s = ctx.solve(tol=1e-8, precision=capi.PREC_FP32_IR, inner_tol=1e-3, maxit=40000)
            out["solve_fp32_ir"] = {k: s[k] for k in ("iters", "outer_iters", "rel_resid", "ms_total")}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
