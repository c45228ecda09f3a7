"""Benchmark: 3D hex8 linear elasticity (96^3 elements, 2,738,019 DOF) on one
MI355X, with the other workloads of the path measured beside it.

One step = one full solve of the resident system to a true relative residual
of 1e-8 with the chosen form (FP32-inner Jacobi-PCG on FP32 3x3-block values
with FP64 iterative refinement).  value = DOF * PCG iterations / second over
the timed steps.  The plain FP64 PCG is measured beside it on the FP64 CSR
and on FP64 3x3 blocks in the same layout, the FP64 assembly is timed separately, the dominant kernel (the SpMV of the inner
solve) is timed with HIP events for the roofline, and the oracle (serial CPU
restatement of the reference path) is timed on a bounded sample.

Also reported in the same JSON line:
  heat    -- transient diffusion 128^3, 100 backward-Euler steps to 1e-8
  packing -- 8 small RVE requests solved one after another and in one pack
  served  -- RVE request stream (the first 1,024 of the seeded 4,096 requests
             by default, ~70 s of arrivals; --served 4096 runs all of them):
             autotuned batch shape, closed-loop capacity against a plain FP64
             server, solves/s and p50/p99 latency at 1e-8, error against FP64
  plastic -- J2 plasticity bar 64^3, 20 increments, algorithmic tangent with
             each Newton correction started from the previous one, and the
             reference's continuum tangent from zero beside it (~110 s)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 96]
                       [--heat-n 128] [--pack-demo 8] [--served 1024] [--plastic-n 64]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "hex8 3D solve: DOF·iter/s & SpMV HBM GB/s; served solves/s + p99 at 1e-8"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(n, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE/WRITE_SIZE scaled by a copy-kernel calibration), for
    the workload size that summary covers; None otherwise."""
    path = os.path.join(ROOT, "profiles", "spmv_hex8_%d.md" % n)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch %s =" % kernel):
            return float(line.split("=")[1])
    return None


def _cpu_info():
    """CPU model, logical CPUs, compiler and the affinity the process runs with."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
            elif line.startswith("Socket(s):") or line.startswith("Thread(s) per core:"):
                info[line.split(":")[0].strip()] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        info["model"] = platform.processor() or platform.machine()
    try:
        info["compiler"] = subprocess.run(["gcc", "--version"], capture_output=True, text=True,
                                          timeout=20).stdout.splitlines()[0]
    except (OSError, subprocess.SubprocessError, IndexError):
        info["compiler"] = None
    return info


def cpu_baseline(sample_n, budget_s):
    """Oracle (serial restatement of SerialAssembler + sparse.cpp PCG, and
    the two-thread ParCG of ParCG.cpp:93-347) at the workload's own size,
    built on this host with -O3 -march=native -ffp-contract=off: pattern and
    assembly once, then as many PCG iterations of each variant as fit the
    budget; the rate is DOF x iterations / s of those iterations."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native"], check=True)
    os.environ["ORACLE_NATIVE"] = "1"
    from betanoox_amd import meshgen
    from oracle import oracle as O
    prob = meshgen.elastic_cube(sample_n)
    t0 = time.perf_counter()
    eq = O.assemble_problem(prob)
    t_asm = time.perf_counter() - t0
    b = eq.rhs()
    rows = {}
    for name, run in (("serial", lambda it: eq.pcg(b, eps=1e-30, maxit=it, stop_rule=1)),
                      ("parcg_2_threads", lambda it: eq.pcg_parcg(b, eps=1e-30, maxit=it))):
        t0 = time.perf_counter()
        run(3)
        per_it = (time.perf_counter() - t0) / 3
        iters = max(5, int(budget_s / max(per_it, 1e-6)))
        t0 = time.perf_counter()
        st, x, it, nrm = run(iters)
        t_pcg = time.perf_counter() - t0
        rows[name] = {"value": prob.n_dof * it / t_pcg, "iterations": it, "seconds": t_pcg,
                      "ms_per_iteration": 1e3 * t_pcg / it}
    info = _cpu_info()
    info["flags"] = "-O3 -march=native -ffp-contract=off -fopenmp (oracle/Makefile native)"
    info["pinning"] = "none (OS scheduling); the ParCG row sets two OpenMP threads itself"
    return {
        "value": rows["serial"]["value"],
        "unit": "DOF*iter/s",
        "cores": 1,
        "kind": "port",
        "sample": "oracle serial Jacobi-PCG (sparse.cpp:114-242 order) on the same generator at %d^3 "
                  "(%d DOF, the workload's own size), %d iterations in %.1f s after its pattern and assembly "
                  "(SerialAssembler order, %.1f s, %.3g elements/s)"
                  % (sample_n, prob.n_dof, rows["serial"]["iterations"], rows["serial"]["seconds"], t_asm,
                     prob.n_el / t_asm),
        "parcg_2_threads": dict(rows["parcg_2_threads"], cores=2,
                                kind="port of ParCGMatrix::solve (ParCG.cpp:93-347)"),
        "assembly_s": t_asm,
        "cpu": info,
    }


def pack_leg(count, reps=3):
    """Packing: `count` small RVE requests (n = 16..24, six load cases each,
    seeded) assembled on their own contexts, solved to 1e-8 one after another
    (bhip_solve_cases per context) and together in one packed solve
    (bhip_pack_solve_cases); best of `reps` for each."""
    from betanoox_amd import capi, meshgen, setup_problem
    sizes = [16 + (i * 8) // max(count - 1, 1) for i in range(count)]
    probs = [meshgen.rve_request(n, 500 + i) for i, n in enumerate(sizes)]
    o = capi.Context.opts(tol=1e-8, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR,
                          inner_tol=1e-3, maxit=40000)
    ctxs = []
    try:
        for p in probs:
            c = capi.Context(0)
            ctxs.append(c)
            setup_problem(c, p)
            c.assemble(capi.ASM_STIFFNESS)
        P = [np.stack(p.extra["cases"]) for p in probs]
        seq = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            for c, pc in zip(ctxs, P):
                c.solve_cases(pc, opts=o)
            seq.append(time.perf_counter() - t0)
        with capi.Pack(0) as pk:
            packed = []
            for _ in range(reps + 1):
                t0 = time.perf_counter()
                _, sts, status = pk.solve_cases(ctxs, P, opts=o)
                packed.append(time.perf_counter() - t0)
        assert all(x == 0 for x in status)
        t_seq, t_pack = min(seq[1:]), min(packed[1:])
        return {"requests": count, "n": sizes, "cases_per_request": 6, "tol": 1e-8,
                "one_after_another_ms": 1e3 * t_seq, "packed_ms": 1e3 * t_pack, "speedup": t_seq / t_pack,
                "inner_iters": [st["iters"] for st in sts]}
    finally:
        for c in ctxs:
            c.close()


def _stream_requests(count, seed):
    """The first `count` requests of the seeded 4,096-request stream (sizes,
    inclusion seeds, exponential inter-arrival gaps of mean 1)."""
    from betanoox_amd import meshgen
    rng = np.random.default_rng(seed)
    total = 4096
    ns = np.rint(np.exp(rng.uniform(np.log(16), np.log(96), total))).astype(int)
    seeds = rng.integers(0, 2 ** 31 - 1, total)
    gaps = rng.exponential(1.0, total)
    reqs = [meshgen.rve_request(int(n), int(sd)) for n, sd in zip(ns[:count], seeds[:count])]
    return reqs, gaps[:count]


def _served_fp64_error(reqs, results, limit=8, max_n=48):
    """Relative displacement errors, worst over the load cases of the first
    `limit` served requests with n <= max_n, against a private plain FP64
    solve of the same request to 1e-12: the served result (FP32-inner,
    residual 1e-8), a private plain FP64 solve to the same residual 1e-8
    (the error the residual test itself allows on these matrices) and a
    private FP32-inner solve to 1e-10."""
    from betanoox_amd import capi, setup_problem
    forms = {"served_fp32_ir_tol_1e-8": None,
             "plain_fp64_tol_1e-8": capi.Context.opts(tol=1e-8, stop_rule=capi.STOP_REL_RESID,
                                                       precision=capi.PREC_FP64, maxit=100000),
             "fp32_ir_tol_1e-10": capi.Context.opts(tol=1e-10, stop_rule=capi.STOP_REL_RESID,
                                                     precision=capi.PREC_FP32_IR, inner_tol=1e-3, maxit=100000)}
    worst = {k: 0.0 for k in forms}
    checked = 0
    o = capi.Context.opts(tol=1e-12, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP64, maxit=100000)
    for rq, X in zip(reqs, results):
        if checked >= limit:
            break
        if round(rq.n_el ** (1 / 3)) > max_n:
            continue
        P = np.stack(rq.extra["cases"])
        with capi.Context(0) as ctx:
            setup_problem(ctx, rq)
            ctx.assemble()
            Xr, _ = ctx.solve_cases(P, opts=o)
            sols = {k: (X if f is None else ctx.solve_cases(P, opts=f)[0]) for k, f in forms.items()}
        for name, Y in sols.items():
            for k in range(Xr.shape[0]):
                worst[name] = max(worst[name], float(np.linalg.norm(Y[k] - Xr[k]) / np.linalg.norm(Xr[k])))
        checked += 1
    return worst, checked


def served_stream(count, rate_frac, probe_count, seed=None, slots=0, max_pack=0):
    """Workload 5 on a prefix of the seeded 4,096-request stream: RVE
    requests (n = round(exp(U(ln 16, ln 96))), 5-40 inclusions, six affine
    load cases each), solved by the resident server to 1e-8.

    1. Batch shape: unless slots and max_pack are given, the autotuner
       (betanoox_amd.autotune.tune_server) times closed-loop runs of a probe
       set (`probe_count` requests of a separate draw, all queued at once)
       for every slots x max_pack candidate and keeps the fastest.
    2. Capacity: the probe set closed-loop in the chosen shape with the
       FP32-inner refined solver, with plain FP64 on the CSR and with plain
       FP64 on FP64 3x3 blocks (same slots and pack),
       at the configured residual 1e-8 and at 1e-10: requests/s and
       DOF*case-iterations/s, ratios.
    3. Stream: the prefix arrives as a Poisson process at `rate_frac` of the
       chosen shape's capacity; solves/s and p50/p99 latency.
    4. Error: served displacements (residual 1e-8) of the first small
       requests against a private FP64 solve to 1e-12, beside a private
       plain FP64 solve to 1e-8 and an FP32-inner solve to 1e-10 of the
       same requests.
    Requests are generated before any clock starts."""
    from betanoox_amd import autotune, capi, meshgen
    seed = meshgen.SEED_BASE + 5 if seed is None else seed
    t0 = time.perf_counter()
    reqs, gaps = _stream_requests(count, seed)
    prng = np.random.default_rng(seed + 1000)
    pns = np.rint(np.exp(prng.uniform(np.log(16), np.log(96), probe_count))).astype(int)
    probe = [meshgen.rve_request(int(n), int(s)) for n, s in zip(pns, prng.integers(0, 2 ** 31 - 1, probe_count))]
    log("served: generated %d stream + %d probe requests in %.1f s" % (count, probe_count, time.perf_counter() - t0))
    ns = [round(r.n_el ** (1 / 3)) for r in reqs]
    out = {"requests": count, "of_stream": 4096, "cases_per_request": 6, "tol": 1e-8,
           "n_min": int(min(ns)), "n_max": int(max(ns)), "dof_total": int(sum(r.n_dof for r in reqs))}
    if slots > 0 and max_pack > 0:
        out["batch_shape"] = {"slots": slots, "max_pack": max_pack, "source": "given"}
    else:
        best, table = autotune.tune_server(probe)
        slots, max_pack = best["slots"], best["max_pack"]
        out["batch_shape"] = {"slots": slots, "max_pack": max_pack, "source": "autotuned (closed-loop probe)",
                              "table": table}
    # closed-loop capacity of every form at the configured residual (1e-8)
    # and at 1e-10 (the residual at which the displacement error, measured
    # below, drops under 1e-8 on these inclusion matrices); DOF*iterations
    # count every load case's iterations (a packed solve advances all its
    # cases in one iteration)
    cap = {}
    for tol in (1e-8, 1e-10):
        for name, prec in (("fp32_ir", capi.PREC_FP32_IR), ("fp64", capi.PREC_FP64),
                           ("fp64_blocked", capi.PREC_FP64_BLOCKED)):
            t, sts = autotune.closed_loop(probe, slots, max_pack, precision=prec, tol=tol)
            dof_it = sum(r.n_dof * st["case_iters"] for r, st in zip(probe, sts))
            key = "%s_tol_%g" % (name, tol)
            cap[key] = {"req_per_s": probe_count / t, "dof_case_iter_per_s": dof_it / t, "wall_s": t,
                        "worst_rel_resid": max(st["rel_resid"] for st in sts)}
            log("served capacity %s: %s" % (key, json.dumps(cap[key])))
    out["capacity"] = cap
    for tol in (1e-8, 1e-10):
        a = cap["fp32_ir_tol_%g" % tol]
        for ref in ("fp64", "fp64_blocked"):
            b = cap["%s_tol_%g" % (ref, tol)]
            out["capacity_ratio_vs_%s_tol_%g" % (ref, tol)] = {
                "req_per_s": a["req_per_s"] / b["req_per_s"],
                "dof_case_iter_per_s": a["dof_case_iter_per_s"] / b["dof_case_iter_per_s"]}
    rate = rate_frac * cap["fp32_ir_tol_1e-08"]["req_per_s"]
    results = []
    with capi.Server(0, max_slots=slots, tol=1e-8, max_pack=max_pack) as srv:
        t0 = time.perf_counter()
        t_next = 0.0
        ids = []
        for k, r in enumerate(reqs):
            t_next += gaps[k] / rate
            dt = t0 + t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(r))
            if (k + 1) % 256 == 0:
                log("served: %d requests submitted" % (k + 1))
        res = []
        for k, i in enumerate(ids):
            X, st = srv.wait(i)
            res.append(st)
            if len(results) < 64:
                results.append(X)
            if st["rel_resid"] > 1e-8:
                raise RuntimeError("served request %d above tolerance: %g" % (k, st["rel_resid"]))
        t_end = time.perf_counter()
    lat = np.array([r["latency_ms"] for r in res])
    span = t_end - t0
    dof_iter = sum(rq.n_dof * st["case_iters"] for rq, st in zip(reqs, res))
    err, checked = _served_fp64_error(reqs, results)
    out.update({
        "arrival": "poisson", "rate_req_per_s": rate, "rate_frac_of_capacity": rate_frac,
        "solves_per_s": count / span, "case_solves_per_s": 6 * count / span,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
        "queue_p50_ms": float(np.percentile([r["queue_ms"] for r in res], 50)),
        "service_p50_ms": float(np.percentile([r["service_ms"] for r in res], 50)),
        "dof_case_iter_per_s": dof_iter / span, "wall_s": span,
        "worst_rel_resid": max(r["rel_resid"] for r in res),
        "pattern_reuse": sum(r["reused_pattern"] for r in res),
        "packed_requests": sum(1 for r in res if r["pack_size"] > 1),
        "mean_pack_size": float(np.mean([r["pack_size"] for r in res])),
        "setup_p50_ms": float(np.percentile([r["setup_ms"] for r in res], 50)),
        "solve_p50_ms": float(np.percentile([r["solve_ms"] for r in res], 50)),
        "rel_error_vs_fp64": {"value": err["served_fp32_ir_tol_1e-8"], **err, "requests_checked": checked,
                              "reference": "private plain FP64 solve of the request to 1e-12"},
    })
    return out


def heat_leg(n, steps, fp64_steps, inner_tol=1e-4):
    """Workload 3: transient diffusion on an n^3 hex8 cube (1 DOF per node),
    backward Euler dt = 1e-3, Mbar assembled once, every step solved to 1e-8
    true relative residual on the device (FP32 inner + FP64 refinement with
    inner reduction `inner_tol`, measured best for this workload), each step's
    initial guess projected on the recent increments (BHIP_HEAT_X0_PROJECTED,
    the chosen form) and, beside it, the previous increment; plain FP64 with
    the projected guess timed on the first `fp64_steps` steps."""
    from betanoox_amd import capi, meshgen, setup_problem
    prob = meshgen.diffusion_cube(n)
    out = {"n": n, "n_dof": prob.n_dof, "steps": steps, "dt": prob.extra["dt"], "tol": 1e-8, "inner_tol": inner_tol}
    with capi.Context(0) as ctx:
        setup_problem(ctx, prob)
        u0 = prob.extra["u0"]
        t0 = time.perf_counter()
        ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[1.0, prob.extra["dt"]], u_state=u0)
        out["assembly_ms"] = 1e3 * (time.perf_counter() - t0)
        chosen, table = ctx.tune_spmv(reps=10)
        out["spmv_tuning"] = {"chosen": table[chosen], "shape0_ms": table[0]["ms"]}
        o32 = capi.Context.opts(tol=1e-8, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR,
                                inner_tol=inner_tol, maxit=40000)
        ctx.heat_run(2, u0, capi.HEAT_X0_PROJECTED, opts=o32)  # warm-up (graphs, tables)
        t0 = time.perf_counter()
        u, st = ctx.heat_run(steps, u0, capi.HEAT_X0_PROJECTED, opts=o32)
        wall = time.perf_counter() - t0
        out.update({"x0": "projected on recent increments", "fp32_ir": st, "wall_s": wall,
                    "steps_per_s": steps / wall, "dof_iter_per_s": prob.n_dof * st["iters"] / wall,
                    "u_max_end": float(np.abs(u).max())})
        t0 = time.perf_counter()
        up, sp = ctx.heat_run(steps, u0, capi.HEAT_X0_PREVIOUS, opts=o32)
        wp = time.perf_counter() - t0
        out["previous_increment_guess"] = {"fp32_ir": sp, "wall_s": wp, "steps_per_s": steps / wp,
                                           "rel_diff_vs_projected": float(np.linalg.norm(up - u) / np.linalg.norm(u))}
        if fp64_steps > 0:
            o64 = capi.Context.opts(tol=1e-8, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP64, maxit=40000)
            t0 = time.perf_counter()
            u64, s64 = ctx.heat_run(fp64_steps, u0, capi.HEAT_X0_PROJECTED, opts=o64)
            w64 = time.perf_counter() - t0
            # error of the chosen form after the same steps, against plain FP64
            # at 1e-8 and at 1e-12
            t0 = time.perf_counter()
            u32, s32f = ctx.heat_run(fp64_steps, u0, capi.HEAT_X0_PROJECTED, opts=o32)
            w32f = time.perf_counter() - t0
            o64t = capi.Context.opts(tol=1e-12, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP64,
                                     maxit=40000)
            u64t, _ = ctx.heat_run(fp64_steps, u0, capi.HEAT_X0_PREVIOUS, opts=o64t)
            out["rel_error_vs_fp64"] = {
                "steps": fp64_steps,
                "fp32_ir_vs_fp64_tol_1e-12": float(np.linalg.norm(u32 - u64t) / np.linalg.norm(u64t)),
                "plain_fp64_tol_1e-8_vs_fp64_tol_1e-12": float(np.linalg.norm(u64 - u64t) / np.linalg.norm(u64t)),
                "fp32_ir_vs_plain_fp64_tol_1e-8": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))}
            out["plain_fp64_first_steps"] = {"steps": fp64_steps, "stats": s64, "wall_s": w64,
                                             "dof_iter_per_s": prob.n_dof * s64["iters"] / w64}
            # the chosen form also builds -K u_s from the FP64 column-run copy
            # of K while plain FP64 reads K's CSR, so this ratio holds a
            # layout gain besides the precision one
            out["fp32_ir_first_steps"] = {"steps": fp64_steps, "stats": s32f, "wall_s": w32f}
            # the same first steps with the same guess in both precisions
            out["speedup_vs_fp64_per_step"] = w64 / w32f
            out["speedup_note"] = ("FP32-inner steps read K for the right-hand side as FP64 column runs, plain "
                                   "FP64 steps as CSR: layout and precision together")
        ms, by = ctx.bench_spmv(fp32=True, reps=20)
        out["spmv_fp32_sell1"] = {"ms": ms, "GBps": by / ms / 1e6, "bytes": by}
    return out


def plastic_leg(n, newton_tol, max_newton, tangent, warm):
    """Workload 4: J2 plasticity bar n^3, 20 prescribed-displacement
    increments to 0.5% strain, Newton with the tangent reassembled every
    iteration until max|residual| < 2*newton_tol, linear solves to 1e-8.
    tangent: the reference's continuum C_ep or the symmetric part of the
    return map's exact derivative (the chosen form); warm: every Newton
    correction starts from the previous one (the chosen form) or from 0."""
    from betanoox_amd import capi, meshgen, setup_problem
    prob = meshgen.plastic_bar(n)
    out = {"n": n, "n_dof": prob.n_dof, "increments": prob.extra["increments"], "newton_tol": newton_tol,
           "tangent": "algorithmic" if tangent == capi.TANGENT_ALGORITHMIC else "continuum (reference)",
           "newton_start": "previous correction" if warm else "zero"}
    with capi.Context(0) as ctx:
        setup_problem(ctx, prob)
        ctx.plastic_init()
        ctx.plastic_set_tangent(tangent)
        ctx.set_load(prob.load)
        o = capi.Context.opts(tol=1e-8, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR,
                              inner_tol=1e-3, maxit=40000, use_x0=1 if warm else 0)
        hist = []
        t0 = time.perf_counter()
        for inc in range(prob.extra["increments"]):
            st = ctx.plastic_increment(prob.prescribed, newton_tol, max_newton, opts=o)
            hist.append(st)
            log("plastic increment %d: %s" % (inc, json.dumps(st)))
        wall = time.perf_counter() - t0
        stress, epsb, _ = ctx.plastic_state(prob.n_el, prob.order ** 3)
    out.update({"wall_s": wall, "newton_iters": sum(h["newton_iters"] for h in hist),
                "pcg_iters": sum(h["pcg_iters"] for h in hist),
                "newton_per_increment": [h["newton_iters"] for h in hist],
                "all_converged": all(h["status"] == 0 for h in hist),
                "plastic_points_end": hist[-1]["plastic_points"],
                "sigma_zz_min_max": [float(stress[:, :, 2].min()), float(stress[:, :, 2].max())],
                "dof_iter_per_s": prob.n_dof * sum(h["pcg_iters"] for h in hist) / wall})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--inner-tol", type=float, default=1e-3)
    ap.add_argument("--cpu-sample", type=int, default=96, help="oracle sample size (the workload's own: 96)")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of PCG iterations per CPU variant")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--served", type=int, default=1024,
                    help="requests of the served stream, a prefix of the 4,096 (0: skip)")
    ap.add_argument("--served-slots", type=int, default=0, help="slots (0: autotuned with --served-pack)")
    ap.add_argument("--served-load", type=float, default=0.7, help="arrival rate / probed capacity")
    ap.add_argument("--served-probe", type=int, default=32)
    ap.add_argument("--served-pack", type=int, default=0, help="requests per packed solve (0: autotuned)")
    ap.add_argument("--pack-demo", type=int, default=8, help="small requests of the packing leg (0: skip)")
    ap.add_argument("--heat-n", type=int, default=128, help="diffusion cube size (0: skip)")
    ap.add_argument("--heat-steps", type=int, default=100)
    ap.add_argument("--heat-fp64-steps", type=int, default=10)
    ap.add_argument("--heat-inner-tol", type=float, default=1e-4)
    ap.add_argument("--plastic-n", type=int, default=64, help="J2 bar size (0: skip; the full workload is 64)")
    ap.add_argument("--plastic-reference-tangent", type=int, default=1,
                    help="also run the reference's continuum tangent (1) or not (0)")
    ap.add_argument("--plastic-tol", type=float, default=1e-9)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (a single solve "
              "would shard by z-slabs with a halo exchange; it is not built); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from betanoox_amd import capi, meshgen, setup_problem

    t_start = time.perf_counter()
    t0 = time.perf_counter()
    prob = meshgen.elastic_cube(args.n)
    log("generated %s: %d nodes, %d elements, %d DOF in %.1f s"
        % (prob.name, prob.n_nodes, prob.n_el, prob.n_dof, time.perf_counter() - t0))
    ctx = capi.Context(0)
    t0 = time.perf_counter()
    setup_problem(ctx, prob)
    t_pattern = time.perf_counter() - t0
    nl, nf = ctx.nnz()
    log("pattern: nnz lower %d, full %d (%.2f s incl. upload)" % (nl, nf, t_pattern))
    # assembly (FP64 integration + scatter), best of 3
    t_asm = []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.assemble(capi.ASM_STIFFNESS)
        t_asm.append(time.perf_counter() - t0)
    ctx.set_load(prob.load)
    ctx.build_rhs()
    log("assembly best %.2f ms" % (1e3 * min(t_asm)))

    result_fp64 = None
    if not args.no_fp64:
        s = ctx.solve(tol=1e-8, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP64, maxit=20000)
        ms64, by64 = ctx.bench_spmv(fp32=False, reps=20)
        x_fp64 = ctx.solution()
        result_fp64 = {
            "iters": s["iters"], "ms": s["ms_total"], "rel_resid": s["rel_resid"],
            "dof_iter_per_s": prob.n_dof * s["iters"] / (s["ms_total"] / 1e3),
            "spmv_ms": ms64, "spmv_GBps": by64 / ms64 / 1e6,
        }
        log("fp64 PCG: %s" % json.dumps(result_fp64))

    # SpMV tile shape: the autotuner times every shape on this matrix and keeps the fastest
    chosen, table = ctx.tune_spmv(reps=10)
    tuning = {"chosen": table[chosen], "shape0_ms": table[0]["ms"], "shapes_timed": len(table)}
    log("spmv tuning: %s" % json.dumps(tuning))

    # plain FP64 Jacobi-PCG on FP64 3x3 blocks in the same SELL layout and tile
    # shape: the FP64 comparison point with the FP32 path's storage
    result_fp64b = None
    if not args.no_fp64:
        s = ctx.solve(tol=1e-8, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP64_BLOCKED, maxit=20000)
        x_fp64b = ctx.solution()
        msb, byb = ctx.bench_spmv(blocked64=True, reps=20)
        result_fp64b = {
            "iters": s["iters"], "ms": s["ms_total"], "rel_resid": s["rel_resid"],
            "dof_iter_per_s": prob.n_dof * s["iters"] / (s["ms_total"] / 1e3),
            "spmv_ms": msb, "spmv_GBps": byb / msb / 1e6, "spmv_bytes": byb,
        }
        log("fp64 blocked PCG: %s" % json.dumps(result_fp64b))

    opts = capi.Context.opts(tol=1e-8, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR,
                             inner_tol=args.inner_tol, maxit=40000)
    for w in range(args.warmup):
        s = ctx.solve(opts)
        log("warmup %d: %s" % (w, json.dumps(s)))
    iters = 0
    outer = 0
    resid = 0.0
    t0 = time.perf_counter()
    for k in range(args.steps):
        s = ctx.solve(opts)
        iters += s["iters"]
        outer += s["outer_iters"]
        resid = max(resid, s["rel_resid"])
        log("step %d: %s" % (k, json.dumps(s)))
    dt = time.perf_counter() - t0
    value = prob.n_dof * iters / dt

    x_fp32 = ctx.solution()
    # error of each form against an FP64 solve to 1e-12 (each result states
    # its error against FP64); the chosen form also at a tighter tolerance
    errors = None
    if not args.no_fp64:
        s_ref = ctx.solve(tol=1e-12, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP64_BLOCKED, maxit=40000)
        x_ref = ctx.solution()
        nref = np.linalg.norm(x_ref)
        s10 = ctx.solve(tol=1e-10, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR,
                        inner_tol=args.inner_tol, maxit=40000)
        x10 = ctx.solution()
        errors = {
            "reference": "plain FP64 PCG on FP64 blocks to 1e-12 (%d iterations, residual %.2e)"
                         % (s_ref["iters"], s_ref["rel_resid"]),
            "fp32_ir_tol_1e-8": float(np.linalg.norm(x_fp32 - x_ref) / nref),
            "plain_fp64_tol_1e-8": float(np.linalg.norm(x_fp64 - x_ref) / nref),
            "fp64_blocked_tol_1e-8": float(np.linalg.norm(x_fp64b - x_ref) / nref),
            "fp32_ir_vs_plain_fp64": float(np.linalg.norm(x_fp32 - x_fp64) / np.linalg.norm(x_fp64)),
            "fp32_ir_tol_1e-10": {"rel_error": float(np.linalg.norm(x10 - x_ref) / nref), "ms": s10["ms_total"],
                                  "iters": s10["iters"], "dof_iter_per_s": prob.n_dof * s10["iters"] /
                                  (s10["ms_total"] / 1e3)},
        }
        log("errors vs FP64: %s" % json.dumps(errors))
    ms32, by32 = ctx.bench_spmv(fp32=True, reps=50)
    achieved = by32 / ms32 / 1e6
    _, stream_gbps = ctx.bench_stream(1 << 30, 20)
    _, read_gbps = ctx.bench_read(2 << 30, 10)
    sh = table[chosen]
    if sh["unroll"] == 0:  # runs of union slots
        kname = "k_sell_spmv_union<1, %d, %d, float>" % (sh["slices_per_block"], sh["nontemporal"])
    else:
        kname = "k_sell_spmv<3, 1, %d, %d, %d, %s, float>" % (sh["slices_per_block"], sh["unroll"], sh["nontemporal"],
                                                              "short" if sh["column_bytes"] == 2 else "int")
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(args.n, kname),
                "kernel": kname + " (FP32 3x3-block sliced-ELL SpMV with p.Ap partials)",
                "bytes_per_launch": by32, "ms_per_launch": ms32}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF*iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 inner / fp64 refinement",
        "data": "synthetic (seeded generator, betanoox_amd/meshgen.py)",
        "config": {
            "workload": "3D linear elasticity, hex8 %d^3 (%d DOF), 2x2x2 Gauss, E log-uniform [1,100], "
                        "nu U[0.2,0.35], jitter 0.1h, z=0 clamped, unit pressure z=1; solve to 1e-8 "
                        "true relative residual" % (args.n, prob.n_dof),
            "n_dof": prob.n_dof, "n_el": prob.n_el, "nnz_full": nf, "nnz_lower": nl,
            "iters_per_solve": iters / args.steps, "outer_per_solve": outer / args.steps,
            "rel_resid_max": resid, "inner_tol": args.inner_tol,
        },
        "roofline": roofline,
        "plain_fp64": result_fp64,
        "fp64_blocked": result_fp64b,
        "assembly": {"ms": 1e3 * min(t_asm), "elements_per_s": prob.n_el / min(t_asm)},
        "spmv_fp32_GBps": achieved,
        "spmv_tuning": tuning,
        "stream_copy_GBps": stream_gbps,
        "stream_read_GBps": read_gbps,
        "spmv_frac_of_stream_read": achieved / read_gbps,
        "rel_error_vs_fp64": errors,
    }
    if result_fp64:
        line["speedup_vs_fp64_dof_iter"] = value / result_fp64["dof_iter_per_s"]
        line["speedup_vs_fp64_time_to_solution"] = result_fp64["ms"] / (1e3 * dt / args.steps)
    if result_fp64b:
        line["speedup_vs_fp64_blocked_dof_iter"] = value / result_fp64b["dof_iter_per_s"]
        line["speedup_vs_fp64_blocked_time_to_solution"] = result_fp64b["ms"] / (1e3 * dt / args.steps)
    ctx.close()
    if args.heat_n > 0:
        line["heat"] = heat_leg(args.heat_n, args.heat_steps, args.heat_fp64_steps, args.heat_inner_tol)
        log("heat: %s" % json.dumps(line["heat"]))
        log("elapsed %.1f s" % (time.perf_counter() - t_start))
    if args.plastic_n > 0:
        line["plastic"] = plastic_leg(args.plastic_n, args.plastic_tol, 300, capi.TANGENT_ALGORITHMIC, True)
        log("plastic: %s" % json.dumps(line["plastic"]))
        if args.plastic_reference_tangent:
            ref = plastic_leg(args.plastic_n, args.plastic_tol, 300, capi.TANGENT_CONTINUUM, False)
            line["plastic"]["reference_tangent"] = ref
            line["plastic"]["speedup_vs_reference_tangent"] = ref["wall_s"] / line["plastic"]["wall_s"]
            log("plastic (reference tangent): %s" % json.dumps(ref))
        log("elapsed %.1f s" % (time.perf_counter() - t_start))
    if args.pack_demo > 0:
        line["packing"] = pack_leg(args.pack_demo)
        log("packing: %s" % json.dumps(line["packing"]))
    if args.served > 0:
        line["served"] = served_stream(args.served, args.served_load, args.served_probe,
                                       slots=args.served_slots, max_pack=args.served_pack)
        log("served: %s" % json.dumps(line["served"]))
        log("elapsed %.1f s" % (time.perf_counter() - t_start))
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample, args.cpu_budget)
    line["bench_wall_s"] = time.perf_counter() - t_start
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
