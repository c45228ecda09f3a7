"""Column-run form of the one-DOF SELL layout through the C ABI.

A scalar matrix (one DOF per node) whose rows hold runs of consecutive
columns, like the 27-column rows of a hex8 diffusion mesh in its natural node
order, is stored as runs of three columns: one 16-bit offset and three values
per run instead of an offset per value.  Checked here:
  - the structured diffusion mesh gets the run form: its FP32 SpMV moves
    fewer bytes than the scalar form (BHIP_SELL_RUNS=0), and both byte counts
    equal a restatement of the two layouts (greedy runs over each node's
    sorted neighbours, slices as wide as their widest node);
  - the refined solve on the run form and on the scalar form both reach the
    plain FP64 solution (the FP32 inner products differ in the order of their
    sums, so the iterates differ in the last bits);
  - the FP64 blocked solve on the run form follows the plain FP64 CSR solve
    (same iteration count, solutions equal to rounding);
  - a randomly renumbered mesh, whose rows have no runs, keeps the scalar form
    (same bytes with and without the switch) and solves to the same answer;
  - every tile shape gives the same bits on the run form (as for the other
    forms, test_gpu_tune).
"""
import os

import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem

pytestmark = pytest.mark.gpu

N_SHAPES = 21


def _permuted(prob, seed):
    """The same mesh with node numbers drawn at random."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(prob.n_nodes)          # old id -> new id
    inv = np.argsort(perm)                        # new id -> old id
    q = meshgen.Problem(prob.name + "_perm", prob.dim, prob.xyz[inv], perm[prob.conn].astype(np.int32),
                        prob.ndof, prob.mat_kind, prob.mat, prob.restrained[inv], prob.prescribed[inv],
                        prob.load[inv], order=prob.order, factor=prob.factor,
                        extra={"u0": prob.extra["u0"][inv], "dt": prob.extra["dt"]})
    return q, inv


def _context(factory, prob, runs):
    old = os.environ.get("BHIP_SELL_RUNS")
    os.environ["BHIP_SELL_RUNS"] = "1" if runs else "0"
    try:
        ctx = setup_problem(factory(), prob)
        ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[1.0, prob.extra["dt"]], u_state=prob.extra["u0"])
        rng = np.random.default_rng(17)
        ctx.set_rhs(np.where(prob.restrained.astype(bool), 0.0, rng.standard_normal(prob.n_dof)))
        # the layout is built by the first FP32 preparation, under the switch
        _, by = ctx.bench_spmv(fp32=True, reps=2)
    finally:
        if old is None:
            del os.environ["BHIP_SELL_RUNS"]
        else:
            os.environ["BHIP_SELL_RUNS"] = old
    return ctx, by


def _layout_bytes(prob):
    """Bytes of one FP32 SpMV in the run form and in the scalar form, from a
    restatement of the layout: a node's sorted neighbour columns covered
    greedily by runs of three consecutive columns; a 64-node slice is as wide
    as its widest node; stored entries + 16-bit offsets + slice pointers + x, y."""
    n = prob.n_nodes
    nbr = [set() for _ in range(n)]
    for el in prob.conn:
        for a in el:
            nbr[a].update(int(b) for b in el)
    runs = np.zeros(n, np.int64)
    deg = np.zeros(n, np.int64)
    for i, cols in enumerate(nbr):
        cols = sorted(cols)
        deg[i] = len(cols)
        k = 0
        while k < len(cols):
            c0 = cols[k]
            runs[i] += 1
            while k < len(cols) and cols[k] <= c0 + 2:
                k += 1
    ns = (n + 63) // 64
    pad = np.zeros(ns * 64, np.int64)
    def slots(per_node):
        pad[:] = 0
        pad[:n] = per_node
        return int(pad.reshape(ns, 64).max(axis=1).sum()) * 64
    tail = (ns + 1) * 4 + 2 * n * 4
    return slots(runs) * 14 + tail, slots(deg) * 6 + tail


def _solve(ctx, precision, tol=1e-8):
    st = ctx.solve(tol=tol, precision=precision, inner_tol=1e-3, maxit=20000)
    return st, ctx.solution()


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("n", [13, 20])
def test_runs_chosen_and_solve(hip_ctx_factory, n):
    print("case runs heat", n, flush=True)
    prob = meshgen.diffusion_cube(n)
    ctx_r, by_r = _context(hip_ctx_factory, prob, True)
    ctx_s, by_s = _context(hip_ctx_factory, prob, False)
    # interior rows: 9 runs x 14 B against 27 entries x 6 B; both layouts as restated
    assert by_r < 0.85 * by_s, (by_r, by_s)
    assert (by_r, by_s) == _layout_bytes(prob)
    st64, x64 = _solve(ctx_s, capi.PREC_FP64, tol=1e-12)
    for ctx in (ctx_r, ctx_s):
        st, x = _solve(ctx, capi.PREC_FP32_IR)
        assert st["rel_resid"] <= 1e-8
        assert _rel(x, x64) <= 1e-7
    # FP64 on the run form: the same recurrence as CSR up to the order of each row's sum
    st_c, x_c = _solve(ctx_s, capi.PREC_FP64)
    st_b, x_b = _solve(ctx_r, capi.PREC_FP64_BLOCKED)
    assert abs(st_b["iters"] - st_c["iters"]) <= 1
    assert _rel(x_b, x_c) <= 1e-8


def test_renumbered_mesh_keeps_scalar_form(hip_ctx_factory):
    prob = meshgen.diffusion_cube(12)
    q, inv = _permuted(prob, 3)
    print("case renumbered heat 12", flush=True)
    ctx_d, by_d = _context(hip_ctx_factory, q, True)
    ctx_s, by_s = _context(hip_ctx_factory, q, False)
    assert by_d == by_s
    st, x = _solve(ctx_d, capi.PREC_FP32_IR)
    assert st["rel_resid"] <= 1e-8
    # the same physical answer as the naturally numbered mesh
    ctx_n, _ = _context(hip_ctx_factory, prob, True)
    rng = np.random.default_rng(17)
    b_q = np.where(q.restrained.astype(bool), 0.0, rng.standard_normal(q.n_dof))
    b_n = np.empty_like(b_q)
    b_n[inv] = b_q
    ctx_n.set_rhs(b_n)
    st_n, x_n = _solve(ctx_n, capi.PREC_FP32_IR)
    assert _rel(x_n[inv], x) <= 1e-7


def test_every_shape_same_bits_on_runs(hip_ctx_factory):
    prob = meshgen.diffusion_cube(13)
    ctx, _ = _context(hip_ctx_factory, prob, True)
    assert 0 <= ctx.spmv_shape() < N_SHAPES
    st0, x0 = _solve(ctx, capi.PREC_FP32_IR)
    for sid in range(N_SHAPES):
        print("shape", sid, flush=True)
        ctx.set_spmv_shape(sid)
        st, x = _solve(ctx, capi.PREC_FP32_IR)
        assert st["iters"] == st0["iters"] and st["outer_iters"] == st0["outer_iters"], sid
        assert np.array_equal(x.view(np.uint64), x0.view(np.uint64)), sid
