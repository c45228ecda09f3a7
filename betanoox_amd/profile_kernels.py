"""Kernel profiling driver for rocprofv3.

Runs a fixed sequence of launches so that kernel traces and PMC counters
can be attributed per kernel, and prints the HIP-event timings of each group.

--workload elastic (default; hex8 elasticity, --n 96):
  reps x plain device copy of 1 GiB       (bhip_bench_stream: k_stream_copy)
  reps x FP32 SELL-64 3x3-block SpMV      (k_sell_spmv<3, 1>)
  reps x FP64 CSR SpMV                    (k_csr_spmv<8,double,double,1>)
  reps x FP64 SELL-64 3x3-block SpMV      (k_sell_spmv<3, 1, ..., double>, BHIP_PREC_FP64_BLOCKED)
  one FP32-inner + FP64-refined solve     (unless --no-solve)
  one reassembly into the solver's layout (k_elem_blocks, k_asm_rows)
--workload cases (RVE request, --n 64, six load cases):
  one six-case solve in the packed solver (k_pk_spmv<6, ...>, k_pk_update<6>, k_pk_pupdate<6>)
--workload heat (diffusion, --n 128, 1 DOF per node):
  reps x FP32 SELL-64 one-DOF SpMV        (k_sell_spmv<1, 1>, column runs)
  reps x FP64 CSR SpMV                    (k_csr_spmv<8,double,double,1>)
  10 backward-Euler steps                 (unless --no-solve)
--workload plastic (J2 bar, --n 48, the first --increments of the
  20-increment path to 0.5% strain): per Newton iteration the FP64 return map
  (k_plastic_update), the tangent assembly (k_elem_blocks, then k_asm_pairs or k_asm_rows of
  ElasHex8<J2>), the device residual (k_plastic_fe, k_plastic_residual) and an
  FP32-inner refined solve to 1e-8
--workload served (RVE requests): --reps requests of the seeded probe draw
  (n = 16..96, six load cases each) queued at once on a 4-slot server
  (pack 8), FP32-inner refined to 1e-8
--workload plastic-kernels (J2 bar, --n 64): the per-Newton-iteration kernels
  without the solve, on the solver's blocked layout: reps x (return map with
  the algorithmic tangent from a plastic trial increment, tangent assembly
  into the blocked layout, device residual)

The SELL SpMV runs in the shape the library's tuned table gives the matrix,
or in --shape ID (see bhip_tune_spmv); the shape is printed.

usage: python -m betanoox_amd.profile_kernels [--workload elastic|cases|heat|plastic|plastic-kernels|served]
                                              [--n N]
                                              [--reps 10]
                                              [--no-solve] [--shape ID] [--increments K]
"""
import argparse
import json
import sys

import numpy as np

from . import capi, meshgen, setup_problem


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("elastic", "cases", "heat", "plastic", "plastic-kernels", "served"),
                    default="elastic")
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--shape", type=int, default=-1)
    ap.add_argument("--increments", type=int, default=6, help="plastic: increments of the 20-increment path run")
    a = ap.parse_args(argv)
    out = {"workload": a.workload}
    if a.workload == "served":
        rng = np.random.default_rng(meshgen.SEED_BASE + 5 + 1000)
        ns = np.rint(np.exp(rng.uniform(np.log(16), np.log(96), a.reps))).astype(int)
        reqs = [meshgen.rve_request(int(n), int(sd)) for n, sd in zip(ns, rng.integers(0, 2 ** 31 - 1, a.reps))]
        with capi.Server(0, max_slots=4, tol=1e-8, max_pack=8) as srv:
            ids = [srv.submit(r) for r in reqs]
            sts = [srv.wait(i)[1] for i in ids]
        out.update({"requests": a.reps, "n": [int(n) for n in ns],
                    "case_iters": int(sum(st["case_iters"] for st in sts)),
                    "worst_rel_resid": max(st["rel_resid"] for st in sts)})
        print(json.dumps(out), flush=True)
        return 0
    with capi.Context(0) as ctx:
        if a.workload == "elastic":
            n = a.n or 96
            prob = meshgen.elastic_cube(n)
            setup_problem(ctx, prob)
            ctx.assemble()
            ctx.set_load(prob.load)
            ctx.build_rhs()
            ms, gbps = ctx.bench_stream(1 << 30, a.reps)
            out["stream_copy"] = {"bytes_read_plus_written": 2 * (1 << 30), "ms": ms, "GBps": gbps}
            if a.shape >= 0:
                ctx.set_spmv_shape(a.shape)
            ms, by = ctx.bench_spmv(fp32=True, reps=a.reps)
            out["sell3_fp32"] = {"bytes": by, "ms": ms, "GBps": by / ms / 1e6, "shape": ctx.spmv_shape()}
            ms, by = ctx.bench_spmv(fp32=False, reps=a.reps)
            out["csr_fp64"] = {"bytes": by, "ms": ms, "GBps": by / ms / 1e6}
            ms, by = ctx.bench_spmv(blocked64=True, reps=a.reps)
            out["sell3_fp64"] = {"bytes": by, "ms": ms, "GBps": by / ms / 1e6}
            if not a.no_solve:
                s = ctx.solve(tol=1e-8, precision=capi.PREC_FP32_IR, inner_tol=1e-3, maxit=40000)
                out["solve_fp32_ir"] = {k: s[k] for k in ("iters", "outer_iters", "rel_resid", "ms_total")}
            # a reassembly on the current blocked layout (built by the SpMV
            # runs above): the gather stage writes the FP32 and FP64 SELL
            # values itself (k_asm_rows, no conversion pass)
            ctx.assemble()
        elif a.workload == "cases":
            n = a.n or 64
            prob = meshgen.rve_request(n, 7)
            setup_problem(ctx, prob)
            ctx.assemble()
            if a.shape >= 0:
                ctx.set_spmv_shape(a.shape)
            P = np.stack(prob.extra["cases"])
            o = capi.Context.opts(tol=1e-8, precision=capi.PREC_FP32_IR, inner_tol=1e-3, maxit=40000)
            X, s = ctx.solve_cases(P, opts=o)
            out["cases_fp32_ir"] = {k: s[k] for k in ("iters", "outer_iters", "rel_resid", "ms_total")}
        elif a.workload == "plastic":
            n = a.n or 48
            prob = meshgen.plastic_bar(n)
            setup_problem(ctx, prob)
            ctx.plastic_init()
            ctx.set_load(prob.load)
            o = capi.Context.opts(tol=1e-8, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR,
                                  inner_tol=1e-3, maxit=40000)
            hist = []
            for inc in range(min(a.increments, prob.extra["increments"])):
                st = ctx.plastic_increment(prob.prescribed, 1e-9, 300, opts=o)
                hist.append(st)
                print("increment %d: %s" % (inc, json.dumps(st)), file=sys.stderr, flush=True)
            out["plastic_increments"] = hist
        elif a.workload == "plastic-kernels":
            n = a.n or 64
            prob = meshgen.plastic_bar(n)
            setup_problem(ctx, prob)
            ctx.plastic_init()
            ctx.plastic_set_tangent(capi.TANGENT_ALGORITHMIC)
            ctx.set_load(prob.load)
            # the solver's blocked layout, as after the first solve of a
            # Newton loop: the tangent assemblies below then gather straight
            # into its FP32 and FP64 blocks (k_asm_rows)
            ctx.assemble()
            ctx.bench_spmv(fp32=True, reps=1)
            ctx.bench_spmv(blocked64=True, reps=1)
            # a uniform axial strain of 0.3 % with lateral contraction: every
            # Gauss point takes the plastic branch
            du = np.zeros(prob.n_dof)
            du[2::3] = 0.003 * prob.xyz[:, 2]
            du[0::3] = -0.0015 * prob.xyz[:, 0]
            du[1::3] = -0.0015 * prob.xyz[:, 1]
            for _ in range(a.reps):
                ctx.plastic_update(du)
                ctx.assemble(capi.ASM_PLASTIC_TANGENT)
                _, mx = ctx.plastic_residual()
            out["max_residual"] = mx
        else:
            n = a.n or 128
            prob = meshgen.diffusion_cube(n)
            setup_problem(ctx, prob)
            u0 = prob.extra["u0"]
            ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[1.0, prob.extra["dt"]], u_state=u0)
            if a.shape >= 0:
                ctx.set_spmv_shape(a.shape)
            ms, by = ctx.bench_spmv(fp32=True, reps=a.reps)
            out["sell1_fp32"] = {"bytes": by, "ms": ms, "GBps": by / ms / 1e6, "shape": ctx.spmv_shape()}
            ms, by = ctx.bench_spmv(fp32=False, reps=a.reps)
            out["csr_fp64"] = {"bytes": by, "ms": ms, "GBps": by / ms / 1e6}
            if not a.no_solve:
                u, s = ctx.heat_run(10, u0, capi.HEAT_X0_PREVIOUS, tol=1e-8, precision=capi.PREC_FP32_IR,
                                    inner_tol=1e-3, maxit=40000)
                out["heat_10_steps"] = s
        out["n"] = n
        out["n_dof"] = prob.n_dof
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
