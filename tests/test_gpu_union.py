"""Union slots of the three-DOF SELL layout through the C ABI.

A slice's slots follow the sorted union of its 64 nodes' neighbour offsets
(zero blocks where a node lacks one), so on a structured hex8 mesh every lane
of a slice gathers the same relative node in a slot and runs of three
consecutive offsets share one gather per lane (the middle value comes from the
next lane).  Checked here:
  - a structured hex8 cube and an RVE request take union slots: the FP32 SpMV
    bytes equal a restatement of the layout (per 64-node slice, the number of
    distinct neighbour offsets; the union runs read values and the offset
    lists, no per-slot columns), and with BHIP_SELL_UNION=0 they equal the
    rank-slot restatement (the widest row of each slice, 16-bit columns);
  - every tile shape, the union runs included, gives the same bits on union
    slots, and a refined solve on union slots equals the solve on rank slots
    (same iteration counts, same values: the slots are summed in the same
    column order, with zero blocks added in between);
  - the FP64 blocked solve on union slots equals the one on rank slots;
  - the packed six-case solve on union slots: large systems (one wave per
    slice) equal the rank-slot solve; small systems (four waves share a slice,
    taking every fourth run) reach the FP64 solution to the solve tolerance;
  - a randomly renumbered mesh, whose slices have wide unions, keeps rank slots
    (same bytes with and without the switch) and solves to the same answer;
  - a reassembly gathers straight into the current layout's blocks (union or
    rank slots) with the bits of a fresh context's conversion.
"""
import dataclasses

import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem

pytestmark = pytest.mark.gpu


def _with_union(monkeypatch, on, fn):
    monkeypatch.setenv("BHIP_SELL_UNION", "1" if on else "0")
    try:
        return fn()
    finally:
        monkeypatch.delenv("BHIP_SELL_UNION")


def _assembled(factory, prob):
    ctx = setup_problem(factory(), prob)
    ctx.assemble(capi.ASM_STIFFNESS)
    ctx.set_load(prob.load)
    ctx.build_rhs()
    # the layout is built by the first FP32 preparation
    _, by = ctx.bench_spmv(fp32=True, reps=2)
    return ctx, by


def _layout_bytes(prob):
    """FP32 SpMV bytes of both slot layouts from the mesh: per 64-node slice,
    rank slots = its widest row, union slots = its distinct neighbour offsets.
    Rank slots: 36 B of values + a 16-bit offset per slot; union runs: the
    values and one 4-byte offset per slot position of a slice (no per-slot
    columns); slice pointers, x and y for both."""
    n = prob.n_nodes
    conn = prob.conn.astype(np.int64)
    k = conn.shape[1]
    a = np.repeat(conn, k, axis=1).ravel()
    b = np.tile(conn, (1, k)).ravel()
    key = np.unique(a * n + b)
    a, b = key // n, key % n
    ns = (n + 63) // 64
    deg = np.bincount(a, minlength=n)
    rank_w = np.zeros(ns, np.int64)
    np.maximum.at(rank_w, np.arange(n) // 64, deg)
    span = 4 * n + 1
    uk = np.unique((a // 64) * span + (b - a + 2 * n))
    union_w = np.bincount(uk // span, minlength=ns)
    tail = (ns + 1) * 4 + 2 * prob.n_dof * 4
    return int(64 * rank_w.sum()) * 38 + tail, int(64 * union_w.sum()) * 36 + int(union_w.sum()) * 4 + tail


def _solve(ctx, tol=1e-8, precision=capi.PREC_FP32_IR):
    st = ctx.solve(tol=tol, precision=precision, inner_tol=1e-3, maxit=40000)
    return st, ctx.solution()


def _permuted(prob, seed):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(prob.n_nodes)
    d = prob.ndof
    dof_new = (perm[:, None] * d + np.arange(d)).ravel()

    def per_dof(a):
        out = np.empty_like(a)
        out[dof_new] = a
        return out
    xyz = np.empty_like(prob.xyz)
    xyz[perm] = prob.xyz
    q = dataclasses.replace(prob, xyz=xyz, conn=perm[prob.conn].astype(np.int32),
                            restrained=per_dof(prob.restrained), prescribed=per_dof(prob.prescribed),
                            load=per_dof(prob.load), extra=dict(prob.extra))
    return q, dof_new


@pytest.mark.parametrize("kind", ["cube13", "rve12"])
def test_union_bytes_and_same_solution(hip_ctx_factory, monkeypatch, kind):
    prob = meshgen.elastic_cube(13) if kind == "cube13" else meshgen.rve_request(12, 3)
    if kind == "rve12":
        prob = dataclasses.replace(prob, load=np.zeros(prob.n_dof))
    print("case", kind, prob.n_dof, flush=True)
    rank_b, union_b = _layout_bytes(prob)
    cu, by_u = _with_union(monkeypatch, True, lambda: _assembled(hip_ctx_factory, prob))
    cr, by_r = _with_union(monkeypatch, False, lambda: _assembled(hip_ctx_factory, prob))
    assert by_u == union_b and by_r == rank_b, (by_u, union_b, by_r, rank_b)
    assert cu.spmv_shape() == 18 and cr.spmv_shape() != 18
    st_u, x_u = _solve(cu)
    st_r, x_r = _solve(cr)
    assert st_u["rel_resid"] <= 1e-8
    assert st_u["iters"] == st_r["iters"] and st_u["outer_iters"] == st_r["outer_iters"]
    assert np.array_equal(x_u, x_r)
    # every shape on the union layout: the same bits
    for sid in range(21):
        print("shape", sid, flush=True)
        cu.set_spmv_shape(sid)
        st, x = _solve(cu)
        assert st["iters"] == st_u["iters"], sid
        assert np.array_equal(x.view(np.uint64), x_u.view(np.uint64)), sid
    # FP64 blocked: the union runs against rank slots
    s64u, x64u = _solve(cu, tol=1e-10, precision=capi.PREC_FP64_BLOCKED)
    s64r, x64r = _solve(cr, tol=1e-10, precision=capi.PREC_FP64_BLOCKED)
    assert s64u["iters"] == s64r["iters"]
    assert np.array_equal(x64u, x64r)


@pytest.mark.parametrize("n", [10, 50])
def test_packed_cases_on_union(hip_ctx_factory, monkeypatch, n):
    """n = 10: a small system (split slices); n = 50: 51^3 nodes, 2073 slices,
    one wave per slice."""
    prob = meshgen.rve_request(n, 5)
    print("case rve", n, prob.n_dof, flush=True)
    P = np.stack(prob.extra["cases"])
    o = capi.Context.opts(tol=1e-10, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR,
                          inner_tol=1e-3, maxit=40000)

    def run(on):
        def go():
            ctx = setup_problem(hip_ctx_factory(), prob)
            ctx.assemble(capi.ASM_STIFFNESS)
            X, st = ctx.solve_cases(P, opts=o)
            _, by = ctx.bench_spmv(fp32=True, reps=1)
            return X, st, by
        return _with_union(monkeypatch, on, go)
    Xu, su, by_u = run(True)
    Xr, sr, by_r = run(False)
    assert su["rel_resid"] <= 1e-10 and sr["rel_resid"] <= 1e-10
    rank_b, union_b = _layout_bytes(prob)
    assert by_u == union_b and by_r == rank_b
    if n >= 50:
        assert su["iters"] == sr["iters"]
        assert np.array_equal(Xu, Xr)
    else:
        err = np.abs(Xu - Xr).max() / np.abs(Xr).max()
        assert err < 1e-8, err


def test_renumbered_mesh_keeps_rank_slots(hip_ctx_factory, monkeypatch):
    prob = meshgen.elastic_cube(12)
    q, dof_new = _permuted(prob, 9)
    print("case renumbered cube12", q.n_dof, flush=True)
    cu, by_u = _with_union(monkeypatch, True, lambda: _assembled(hip_ctx_factory, q))
    cr, by_r = _with_union(monkeypatch, False, lambda: _assembled(hip_ctx_factory, q))
    assert by_u == by_r
    assert cu.spmv_shape() != 18
    st_q, x_q = _solve(cu, tol=1e-10)
    c0, _ = _assembled(hip_ctx_factory, prob)
    st_0, x_0 = _solve(c0, tol=1e-10)
    err = np.abs(x_q[dof_new] - x_0).max() / np.abs(x_0).max()
    assert err < 1e-7, err


@pytest.mark.parametrize("kind", ["cube13", "rve12", "renumbered12"])
def test_reassembly_into_the_layout(hip_ctx_factory, kind):
    """A reassembly on a current layout gathers straight into its FP32 and
    FP64 blocks, slot by slot (union slots: each slot's offset looked up in
    the node's neighbour list; rank slots on the renumbered mesh: the node's
    k-th neighbour).  Solves after it equal, bit for bit, the solves of a
    fresh context, whose first assembly writes the CSR and converts it."""
    if kind == "cube13":
        a = meshgen.elastic_cube(13)
    elif kind == "rve12":
        a = meshgen.rve_request(12, 3)
    else:
        a, _ = _permuted(meshgen.elastic_cube(12), 4)
    b = dataclasses.replace(a, mat=a.mat.copy())
    b.mat[0, 0] *= 3.0  # the first material group's modulus
    print("case reassembly", kind, a.n_dof, flush=True)
    ctx, _ = _assembled(hip_ctx_factory, a)
    assert (ctx.spmv_shape() == 18) == (kind != "renumbered12")
    _, xa = _solve(ctx)
    _, xa64 = _solve(ctx, tol=1e-10, precision=capi.PREC_FP64_BLOCKED)
    ctx.set_materials(b.mat_kind, b.mat, b.extra.get("group"))
    ctx.assemble(capi.ASM_STIFFNESS)
    ctx.build_rhs()
    _, xb = _solve(ctx)
    _, xb64 = _solve(ctx, tol=1e-10, precision=capi.PREC_FP64_BLOCKED)
    fb, _ = _assembled(hip_ctx_factory, b)
    _, fxb = _solve(fb)
    _, fxb64 = _solve(fb, tol=1e-10, precision=capi.PREC_FP64_BLOCKED)
    assert np.array_equal(xb, fxb)
    assert np.array_equal(xb64, fxb64)
    assert not np.array_equal(xa, xb) and not np.array_equal(xa64, xb64)
