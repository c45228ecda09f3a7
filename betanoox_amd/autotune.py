"""Autotuner of the SpMV tile shapes and of the server's batch shape.

Batch shape (tune_server): the served solver's slot count (requests on the
device at once) and pack size (small requests solved together in one packed
batched PCG) are picked by timing closed-loop runs of a probe set of
requests, every combination of the candidate values, and keeping the one with
the highest capacity (requests per second with the whole probe queued).

SpMV tile shapes, two levels:
  * at run time, bhip_tune_spmv (Context.tune_spmv) times every SELL tile
    shape inside inner PCG iterations on a context's own matrix and keeps the
    shape with the fastest iteration;
  * this module sweeps that tuner over the seeded workload generators (1, 2
    and 3 DOFs per node, a range of mesh sizes) and, for the packed SpMV of
    the multi-case and served path, times six-case solves with streaming
    loads on and off (BHIP_PACK_STREAMING; the packed solver streams when its
    systems' SELL arrays exceed 240 MB).  --emit-header turns a sweep into csrc/tuned_shapes.hpp, the table a
    context uses when nobody tuned it (rebuild the library afterwards), and
    --emit-md into a short summary for profiles/.

Shapes never change results (every shape sums in the same order), so the
table only decides speed.

usage:
  python -m betanoox_amd.autotune --sweep OUT.json [--quick]     (GPU)
  python -m betanoox_amd.autotune --emit-header OUT.json          (CPU)
  python -m betanoox_amd.autotune --emit-md OUT.json MD           (CPU)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "csrc", "tuned_shapes.hpp")

# (generator, B, sizes): sizes span small matrices that sit in the 256 MB
# last-level cache to the full workload sizes
SWEEP = [
    ("elastic_cube", 3, [16, 24, 32, 40, 48, 56, 64, 72, 80, 96, 112]),
    ("diffusion_cube", 1, [32, 48, 64, 80, 96, 112, 128, 160]),
    ("plane_stress_quad", 2, [128, 256, 512, 768, 1024, 1536, 2048]),
]
QUICK = [
    ("elastic_cube", 3, [24, 48, 96]),
    ("diffusion_cube", 1, [64, 128]),
    ("plane_stress_quad", 2, [256, 1024]),
]
RVE_SIZES = [16, 24, 32, 48, 64, 80, 96]
TIE = 0.01                          # shapes within 1% + 0.2 us of the best count as equal
TIE_US = 0.2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def closed_loop(reqs, slots, max_pack, precision=None, tol=1e-8):
    """Serve every request of `reqs` queued at once on a fresh server;
    returns (seconds, per-request stats)."""
    from . import capi
    prec = capi.PREC_FP32_IR if precision is None else precision
    with capi.Server(0, max_slots=slots, tol=tol, max_pack=max_pack, precision=prec) as srv:
        t0 = time.perf_counter()
        ids = [srv.submit(r) for r in reqs]
        sts = [srv.wait(i)[1] for i in ids]
        return time.perf_counter() - t0, sts


def tune_server(probe, slots_list=(2, 4, 8), pack_list=(1, 4, 8, 16), reps=1):
    """Batch shape of the served solver: closed-loop capacity of the probe
    requests for every (slots, max_pack); returns (best, table), best =
    {"slots", "max_pack", "req_per_s"}.  A warm-up run on the first
    combination comes first (kernel loading, host allocations).  Capacities
    within 2% of the best tie; the tie goes to the slot count nearest 4."""
    closed_loop(probe[: max(2, len(probe) // 4)], slots_list[0], pack_list[0])
    table = []
    for slots in slots_list:
        for mp in pack_list:
            best_t = min(closed_loop(probe, slots, mp)[0] for _ in range(reps))
            table.append({"slots": slots, "max_pack": mp, "req_per_s": len(probe) / best_t})
            log("server tuning: slots %d, max_pack %d: %.2f req/s" % (slots, mp, table[-1]["req_per_s"]))
    # capacities within 2% count as equal (run-to-run spread of a 32-request
    # probe); among them the slot count nearest 4, then the fewer slots, then
    # the smallest pack win.  Measured on the seeded RVE stream at 70% of
    # capacity (pack 1): 2 slots p50 48.7 ms / p99 636 ms (a large request
    # holds one of the two workers and small ones queue behind it), 4 slots
    # 40.9 / 565 ms, 8 slots 43.2 / 693 ms (more requests share the GPU).
    top = max(r["req_per_s"] for r in table)
    near = [r for r in table if r["req_per_s"] >= 0.98 * top]
    best = min(near, key=lambda r: (abs(math.log2(r["slots"] / 4.0)), r["slots"], r["max_pack"]))
    return best, table


def _assembled(ctx, gen, n):
    from . import capi, meshgen, setup_problem
    prob = getattr(meshgen, gen)(n)
    setup_problem(ctx, prob)
    if gen == "diffusion_cube":
        ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[1.0, prob.extra["dt"]], u_state=prob.extra["u0"])
    else:
        ctx.assemble(capi.ASM_STIFFNESS)
    return prob


def sweep(quick=False, reps=10):
    from . import capi, meshgen, setup_problem
    out = {"device": "gfx950", "reps": reps, "single": [], "multi": []}
    for gen, b, sizes in (QUICK if quick else SWEEP):
        for n in sizes:
            with capi.Context(0) as ctx:
                prob = _assembled(ctx, gen, n)
                chosen, table = ctx.tune_spmv(reps=reps)
                n_slices = (prob.n_nodes + 63) // 64
                rec = {"gen": gen, "b": b, "n": n, "n_nodes": prob.n_nodes, "n_slices": n_slices,
                       "chosen": chosen, "table": table}
                out["single"].append(rec)
                best = table[chosen]
                log("%s %d (B=%d, %d slices): shape %d, iteration %.1f us, SpMV %.1f us %.0f GB/s; shape 0 %.1f / %.1f us"
                    % (gen, n, b, n_slices, chosen, best["iter_ms"] * 1e3, best["ms"] * 1e3, best["gbps"],
                       table[0]["iter_ms"] * 1e3, table[0]["ms"] * 1e3))
    o = capi.Context.opts(tol=1e-8, precision=capi.PREC_FP32_IR, inner_tol=1e-3, maxit=40000)
    for n in ([24, 64] if quick else RVE_SIZES):
        prob = meshgen.rve_request(n, 7)
        P = np.stack(prob.extra["cases"])
        rec = {"gen": "rve_request", "b": 3, "n": n, "n_nodes": prob.n_nodes,
               "n_slices": (prob.n_nodes + 63) // 64, "ms": {}}
        with capi.Context(0) as ctx:
            setup_problem(ctx, prob)
            ctx.assemble()
            for label, flag in (("plain", "0"), ("stream", "1")):
                os.environ["BHIP_PACK_STREAMING"] = flag
                ctx.solve_cases(P, opts=o)          # warm-up (graph capture)
                best = 1e300
                for _ in range(3):
                    X, s = ctx.solve_cases(P, opts=o)
                    best = min(best, s["ms_total"])
                rec["ms"][label] = best
                rec["iters"] = s["iters"]
            os.environ.pop("BHIP_PACK_STREAMING", None)
        out["multi"].append(rec)
        log("rve %d: six cases plain %.2f ms, streaming %.2f ms" % (n, rec["ms"]["plain"], rec["ms"]["stream"]))
    return out


def _rows(recs):
    """Contiguous size ranges with one shape each: starting from the default
    shape 0, keep the current shape while it stays within TIE (relative) plus
    TIE_US (absolute, launch-size noise) of the best at the next size, else
    switch to that size's fastest; a boundary sits at the geometric mean of
    the slice counts either side."""
    recs = sorted(recs, key=lambda r: r["n_slices"])
    rows = []
    cur = 0
    for r in recs:
        ms = {t["id"]: t["iter_ms"] for t in r["table"]}
        best = min(ms.values())
        if ms[cur] > best * (1 + TIE) + TIE_US * 1e-3:
            cur = r["chosen"]
            rows.append({"shape": cur, "first": r["n_slices"], "last": r["n_slices"]})
        elif not rows:
            rows.append({"shape": cur, "first": r["n_slices"], "last": r["n_slices"]})
        else:
            rows[-1]["last"] = r["n_slices"]
    out = []
    for i, row in enumerate(rows):
        if i + 1 < len(rows):
            hi = int(math.sqrt(row["last"] * rows[i + 1]["first"]))
        else:
            hi = 1 << 40
        out.append((hi, row["shape"]))
    return out


def emit_header(data, path=HEADER):
    lines = [
        "// tuned_shapes.hpp -- default tile shapes of the FP32 SELL SpMV on gfx950.",
        "// Written by `python -m betanoox_amd.autotune --emit-header` from a sweep of",
        "// bhip_tune_spmv over the seeded workload generators; edit by re-running it.",
        "// A row applies to matrices with B DOFs per node and at most max_slices 64-node",
        "// slices; the first matching row wins.  Shape ids are those of bhip_tune_spmv.",
        "#pragma once",
        "",
        "struct TunedSellRow {",
        "    int b;",
        "    long long max_slices;",
        "    int shape;",
        "};",
        "",
        "constexpr TunedSellRow kTunedSell[] = {",
    ]
    for b in (3, 2, 1):
        recs = [r for r in data["single"] if r["b"] == b]
        if not recs:
            continue
        sizes = ", ".join("%d" % r["n"] for r in sorted(recs, key=lambda r: r["n"]))
        lines.append("    // B = %d, measured at n = %s (%s)" % (b, sizes, recs[0]["gen"]))
        for hi, shape in _rows(recs):
            lim = "1LL << 40" if hi >= (1 << 40) else "%d" % hi
            lines.append("    {%d, %s, %d}," % (b, lim, shape))
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def emit_md(data, path):
    L = ["# SpMV tile-shape sweep (bhip_tune_spmv), gfx950", "",
         "Command (GPU): `python -m betanoox_amd.autotune --sweep autotune.json`; "
         "summary: `python -m betanoox_amd.autotune --emit-md autotune.json <this file>`.", "",
         "Each shape: best of three interleaved rounds of %d inner PCG iterations (HIP events around "
         "the SpMV and around the whole iteration); the fastest iteration picks the shape. GB/s = "
         "algorithmic bytes of the stored form (B x B blocks, or column runs of three for the one-DOF "
         "diffusion rows: values + 16-bit offsets + slice pointers + x + y) / SpMV time." % data["reps"], "",
         "| generator | n | B | slices | fastest shape | iteration us | SpMV us | GB/s | shape 0 iteration / SpMV us |",
         "|---|---|---|---|---|---|---|---|---|"]
    for r in data["single"]:
        t = r["table"][r["chosen"]]
        t0 = r["table"][0]
        desc = "%d (spb %d, unroll %d, nt %d)" % (t["id"], t["slices_per_block"], t["unroll"], t["nontemporal"])
        L.append("| %s | %d | %d | %d | %s | %.1f | %.1f | %.0f | %.1f / %.1f |"
                 % (r["gen"], r["n"], r["b"], r["n_slices"], desc, t["iter_ms"] * 1e3, t["ms"] * 1e3, t["gbps"],
                    t0["iter_ms"] * 1e3, t0["ms"] * 1e3))
    if data["multi"]:
        L += ["", "Six-case RVE solves (bhip_solve_cases, FP32 inner + FP64 refinement to 1e-8), "
              "multi-case SpMV with plain or streaming loads (best of 3 solves):", "",
              "| n | slices | inner iterations | plain ms | streaming ms |", "|---|---|---|---|---|"]
        for r in data["multi"]:
            L.append("| %d | %d | %d | %.2f | %.2f |" % (r["n"], r["n_slices"], r["iters"], r["ms"]["plain"],
                                                       r["ms"]["stream"]))
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
    return path


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sweep", metavar="OUT")
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--emit-header", metavar="JSON")
    ap.add_argument("--emit-md", nargs=2, metavar=("JSON", "MD"))
    a = ap.parse_args(argv)
    if a.sweep:
        t0 = time.perf_counter()
        data = sweep(a.quick, a.reps)
        data["wall_s"] = time.perf_counter() - t0
        with open(a.sweep, "w") as f:
            json.dump(data, f)
        log("sweep written to %s (%.1f s)" % (a.sweep, data["wall_s"]))
    if a.emit_header:
        print(emit_header(json.load(open(a.emit_header))))
    if a.emit_md:
        print(emit_md(json.load(open(a.emit_md[0])), a.emit_md[1]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
