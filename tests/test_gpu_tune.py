"""SpMV tile-shape autotuner (bhip_tune_spmv) through the C ABI.

A tile shape changes the speed of the FP32 blocked SpMV, never its results:
every shape sums each node row and the p.Ap partials in the same order (the
union runs, shapes 18-20, too: on other layouts they run as unroll 4), so a
solve gives the same bits whichever shape runs (checked bit for bit here on
meshes whose slice count is not a multiple of the workgroup sizes, for 1, 2
and 3 DOFs per node).  The tuner reports every shape, keeps the fastest, and a
solve after tuning equals the solve before it bit for bit.  The tuner keeps
the shape whose inner PCG iteration (SpMV + fused vector kernels) is fastest.
"""
import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem

pytestmark = pytest.mark.gpu

N_SHAPES = 21


def _problem(kind):
    if kind == "hex11":        # 1728 nodes: 27 slices
        return meshgen.elastic_cube(11), capi.ASM_STIFFNESS, None
    if kind == "quad40":       # 1681 nodes, 2 DOFs per node: 27 slices
        return meshgen.plane_stress_quad(40), capi.ASM_STIFFNESS, None
    if kind == "heat13":       # 2744 nodes, 1 DOF per node: 43 slices
        p = meshgen.diffusion_cube(13)
        return p, capi.ASM_HEAT_TRANSIENT, p
    raise KeyError(kind)


def _assembled(factory, kind):
    prob, what, heat = _problem(kind)
    print("case", kind, prob.n_dof, flush=True)
    ctx = setup_problem(factory(), prob)
    if heat is not None:
        ctx.assemble(what, coef=[1.0, heat.extra["dt"]], u_state=heat.extra["u0"])
        rng = np.random.default_rng(7)
        b = rng.standard_normal(prob.n_dof)
        ctx.set_rhs(b)
    else:
        ctx.assemble(what)
        ctx.set_load(prob.load)
        ctx.build_rhs()
    return prob, ctx


def _solve(ctx):
    st = ctx.solve(tol=1e-8, precision=capi.PREC_FP32_IR, inner_tol=1e-3, maxit=20000)
    return st, ctx.solution()


@pytest.mark.parametrize("kind", ["hex11", "quad40", "heat13"])
def test_every_shape_same_bits(hip_ctx_factory, kind):
    prob, ctx = _assembled(hip_ctx_factory, kind)
    st0, x0 = _solve(ctx)
    assert st0["rel_resid"] <= 1e-8
    # default shape once the layout exists: the union runs on hex8 (union
    # slots), the tuned table otherwise
    default = ctx.spmv_shape()
    assert default == (18 if kind == "hex11" else 0)
    for sid in range(N_SHAPES):
        if sid == default:
            continue
        print("shape", sid, flush=True)
        ctx.set_spmv_shape(sid)
        assert ctx.spmv_shape() == sid
        st, x = _solve(ctx)
        assert st["iters"] == st0["iters"] and st["outer_iters"] == st0["outer_iters"], sid
        assert np.array_equal(x.view(np.uint64), x0.view(np.uint64)), sid


def test_tune_keeps_fastest_and_bits(hip_ctx_factory):
    prob, ctx = _assembled(hip_ctx_factory, "hex11")
    st0, x0 = _solve(ctx)
    chosen, table = ctx.tune_spmv(reps=5)
    assert len(table) == N_SHAPES
    assert [t["id"] for t in table] == list(range(N_SHAPES))
    ms = np.array([t["ms"] for t in table])
    it_ms = np.array([t["iter_ms"] for t in table])
    assert np.all(ms > 0) and np.all(np.isfinite(ms))
    assert np.all(it_ms > ms) and np.all(np.isfinite(it_ms))
    assert chosen == int(np.argmin(it_ms))
    assert all(t["column_bytes"] == 2 for t in table)  # 16-bit offsets on this mesh
    assert ctx.spmv_shape() == chosen
    st, x = _solve(ctx)
    assert st["iters"] == st0["iters"]
    assert np.array_equal(x.view(np.uint64), x0.view(np.uint64))


def test_shape_argument_errors(hip_ctx_factory):
    prob, ctx = _assembled(hip_ctx_factory, "hex11")
    with pytest.raises(capi.BhipError):
        ctx.set_spmv_shape(N_SHAPES)
    with pytest.raises(capi.BhipError):
        ctx.set_spmv_shape(-1)
    assert ctx.spmv_shape() == 0  # unchanged (no layout built yet)


def _permuted(prob, seed):
    """The same mesh with its nodes renumbered at random (wide node bandwidth)."""
    import dataclasses
    rng = np.random.default_rng(seed)
    perm = rng.permutation(prob.n_nodes)          # old node i -> new node perm[i]
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


def test_narrow_and_wide_columns(hip_ctx_factory, monkeypatch):
    """16-bit column offsets and int32 columns give the same bits; a mesh whose
    node numbering spans more than 32767 nodes between neighbours runs on the
    int32 columns and agrees with the natural numbering."""
    prob, _, _ = _problem("hex11")
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.assemble(capi.ASM_STIFFNESS)
    ctx.set_load(prob.load)
    ctx.build_rhs()
    st_n, x_n = _solve(ctx)
    cases = np.stack([prob.prescribed] * 2)
    o = capi.Context.opts(tol=1e-8, precision=capi.PREC_FP32_IR, inner_tol=1e-3, maxit=20000)
    Xc_n, _ = ctx.solve_cases(cases, opts=o)
    monkeypatch.setenv("BHIP_SELL_WIDE_COLUMNS", "1")
    ctx_w = setup_problem(hip_ctx_factory(), prob)
    ctx_w.assemble(capi.ASM_STIFFNESS)
    ctx_w.set_load(prob.load)
    ctx_w.build_rhs()
    st_w, x_w = _solve(ctx_w)
    Xc_w, _ = ctx_w.solve_cases(cases, opts=o)
    _, by_n = ctx.bench_spmv(fp32=True, reps=1)
    _, by_w = ctx_w.bench_spmv(fp32=True, reps=1)
    monkeypatch.delenv("BHIP_SELL_WIDE_COLUMNS")
    assert by_w > by_n  # 4-byte against 2-byte columns
    assert st_w["iters"] == st_n["iters"]
    assert np.array_equal(x_w.view(np.uint64), x_n.view(np.uint64))
    assert np.array_equal(Xc_w, Xc_n)
    # natural (narrow) against randomly renumbered (wide) numbering, 33^3 nodes
    big = meshgen.elastic_cube(32)
    perm_prob, dof_new = _permuted(big, 5)
    res = []
    for p in (big, perm_prob):
        c = setup_problem(hip_ctx_factory(), p)
        c.assemble(capi.ASM_STIFFNESS)
        c.set_load(p.load)
        c.build_rhs()
        st = c.solve(tol=1e-10, precision=capi.PREC_FP32_IR, inner_tol=1e-3, maxit=40000)
        assert st["rel_resid"] <= 1e-10
        res.append(c.solution())
    x_nat, x_perm = res
    err = np.abs(x_perm[dof_new] - x_nat).max() / np.abs(x_nat).max()
    assert err < 1e-7, err
