"""Plain FP64 Jacobi-PCG on FP64 node blocks (BHIP_PREC_FP64_BLOCKED) through
the C ABI.

The recurrence is the reference's FP64 Jacobi-PCG (LargeMatrix::solve,
sparse.cpp:144-240); only the matrix storage differs from BHIP_PREC_FP64: the
assembled FP64 values held as B x B node blocks in the SELL-64 layout of the
FP32 inner solve.  The blocked SpMV adds a row's terms in block order, not
in CSR order, so the bits differ from BHIP_PREC_FP64 while the solution does
not beyond the tolerance.  Checks (tolerances per check):
  - against the oracle's FP64 PCG solved to 1e-12 (relative residual): true
    relative residual <= 1e-11, displacement error <= 1e-8 relative,
    iteration count within 2 of the oracle's, for 2 and 3 DOFs per node and
    nonzero prescribed values
  - 1 DOF per node (transient diffusion matrix): solution within 1e-9
    relative of the FP64 CSR solve, both to 1e-12
  - the first 1, 2 and 5 iterates equal those of the FP64 CSR solve within
    1e-12 relative (they carry the SpMV through alpha and the residual)
  - every tile shape gives the same bits, and so do int32 block columns
    against 16-bit offsets; the same solve twice gives the same bits
  - several load cases (solved in turn) equal single solves bit for bit
  - a matrix without a mesh behind it -> BHIP_ERR_ARG
  - bench_spmv reports the algorithmic bytes of the FP64 blocked form
"""
import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N_SHAPES = 21


def _problem(kind):
    if kind == "quad16":
        return meshgen.plane_stress_quad(16)
    if kind == "hex6":
        return meshgen.elastic_cube(6)
    if kind == "hex5_rollers":
        p = meshgen.plastic_bar(5)
        p.mat_kind = capi.MAT_ISO_ELASTIC
        p.mat = np.array([[200.0, 0.3]])
        p.prescribed = p.prescribed * 100.0
        return p
    if kind == "hex11":
        return meshgen.elastic_cube(11)
    raise KeyError(kind)


def _ready(factory, prob):
    ctx = setup_problem(factory(), prob)
    ctx.assemble()
    ctx.set_load(prob.load)
    ctx.build_rhs()
    return ctx


def _o64b(tol, stop_rule=capi.STOP_REL_RESID, maxit=20000):
    return capi.Context.opts(tol=tol, stop_rule=stop_rule, precision=capi.PREC_FP64_BLOCKED, maxit=maxit)


@pytest.mark.parametrize("kind", ["quad16", "hex6", "hex5_rollers"])
def test_against_oracle(hip_ctx_factory, kind):
    prob = _problem(kind)
    print("case", kind, prob.n_dof, flush=True)
    ctx = _ready(hip_ctx_factory, prob)
    eq = O.assemble_problem(prob)
    b = eq.rhs()
    st, x_o, it_o, _ = eq.pcg(b, eps=1e-12, stop_rule=1)
    assert st == 0
    s = ctx.solve(_o64b(1e-12))
    x_g = ctx.solution()
    assert s["rel_resid"] <= 1e-11
    assert np.linalg.norm(x_g - x_o) <= 1e-8 * np.linalg.norm(x_o)
    assert abs(s["iters"] - it_o) <= 2
    # the reference's stopping rule sqrt(|r|^2/|x|^2) < eps (sparse.cpp:223)
    st, _, it_o2, _ = eq.pcg(b, eps=1e-8, stop_rule=0)
    s2 = ctx.solve(_o64b(1e-8, capi.STOP_REF_R_OVER_X))
    assert abs(s2["iters"] - it_o2) <= 2
    x3 = ctx.solution()
    ctx.solve(_o64b(1e-8, capi.STOP_REF_R_OVER_X))
    assert np.array_equal(x3.view(np.uint64), ctx.solution().view(np.uint64))


@pytest.mark.parametrize("k", [1, 2, 5])
def test_early_iterates_match_csr(hip_ctx_factory, k):
    """x1 = alpha z0 with alpha = r0.z0 / z0.K z0; later iterates also carry
    K p through the residual.  FP64 blocks and FP64 CSR differ by rounding
    only (1e-12 relative)."""
    prob = _problem("hex6")
    ctx = _ready(hip_ctx_factory, prob)
    s64 = ctx.solve(tol=1e-30, maxit=k, precision=capi.PREC_FP64, allow_not_converged=True)
    x_csr = ctx.solution()
    sb = ctx.solve(_o64b(1e-30, maxit=k), allow_not_converged=True)
    x_blk = ctx.solution()
    assert s64["iters"] == sb["iters"] == k
    assert np.abs(x_blk - x_csr).max() <= 1e-12 * np.abs(x_csr).max()


def test_one_dof_per_node(hip_ctx_factory):
    p = meshgen.diffusion_cube(13)
    ctx = setup_problem(hip_ctx_factory(), p)
    ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[1.0, p.extra["dt"]], u_state=p.extra["u0"])
    b = np.random.default_rng(5).standard_normal(p.n_dof)
    ctx.set_rhs(b)
    ctx.solve(tol=1e-12, precision=capi.PREC_FP64, maxit=20000)
    x_csr = ctx.solution()
    s = ctx.solve(_o64b(1e-12))
    assert s["rel_resid"] <= 1e-11
    x = ctx.solution()
    assert np.linalg.norm(x - x_csr) <= 1e-9 * np.linalg.norm(x_csr)


def test_every_shape_same_bits(hip_ctx_factory):
    ctx = _ready(hip_ctx_factory, _problem("hex11"))
    s0 = ctx.solve(_o64b(1e-10))
    x0 = ctx.solution()
    for sid in range(N_SHAPES):
        print("shape", sid, flush=True)
        ctx.set_spmv_shape(sid)
        s = ctx.solve(_o64b(1e-10))
        assert s["iters"] == s0["iters"], sid
        assert np.array_equal(ctx.solution().view(np.uint64), x0.view(np.uint64)), sid


def test_wide_columns_same_bits(hip_ctx_factory, monkeypatch):
    """int32 block columns (the form for meshes whose neighbours lie more than
    32767 nodes apart) give the same bits as 16-bit offsets."""
    prob = _problem("hex11")
    ctx = _ready(hip_ctx_factory, prob)
    s_n = ctx.solve(_o64b(1e-10))
    x_n = ctx.solution()
    _, by_n = ctx.bench_spmv(blocked64=True, reps=1)
    monkeypatch.setenv("BHIP_SELL_WIDE_COLUMNS", "1")
    ctx_w = _ready(hip_ctx_factory, prob)
    s_w = ctx_w.solve(_o64b(1e-10))
    _, by_w = ctx_w.bench_spmv(blocked64=True, reps=1)
    monkeypatch.delenv("BHIP_SELL_WIDE_COLUMNS")
    assert by_w > by_n  # 4-byte against 2-byte columns
    assert s_w["iters"] == s_n["iters"]
    assert np.array_equal(ctx_w.solution().view(np.uint64), x_n.view(np.uint64))


def test_cases_in_turn(hip_ctx_factory):
    prob = meshgen.rve_request(6, 9)
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.assemble(capi.ASM_STIFFNESS)
    P = np.stack(prob.extra["cases"])[:3]
    o = _o64b(1e-12)
    X, st = ctx.solve_cases(P, opts=o)
    for k in range(3):
        ctx.set_prescribed(P[k])
        ctx.assemble(capi.ASM_STIFFNESS)
        ctx.set_load(prob.load)
        ctx.build_rhs()
        ctx.solve(o)
        assert np.array_equal(X[k].view(np.uint64), ctx.solution().view(np.uint64)), k


def test_matrix_without_mesh_rejected(hip_ctx_factory):
    n = 6
    A = np.diag(np.full(n, 4.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    ia = np.concatenate([[0], np.cumsum((A != 0).sum(1))]).astype(np.int32)
    ja = np.nonzero(A)[1].astype(np.int32)
    ctx = hip_ctx_factory()
    ctx.set_csr(ia, ja, A[A != 0])
    ctx.set_rhs(np.ones(n))
    with pytest.raises(capi.BhipError) as ei:
        ctx.solve(_o64b(1e-10))
    assert ei.value.status == capi.ERR_ARG
    s = ctx.solve(tol=1e-10, precision=capi.PREC_FP64)
    assert s["rel_resid"] <= 1e-10


def test_bench_bytes(hip_ctx_factory):
    ctx = _ready(hip_ctx_factory, _problem("hex11"))
    ms32, by32 = ctx.bench_spmv(fp32=True, reps=3)
    ms64, by64 = ctx.bench_spmv(blocked64=True, reps=3)
    assert ms32 > 0 and ms64 > 0
    n_dof = ctx.n_dof
    # the hex mesh has union slots and runs in the union shape: per padded
    # block 36 B of FP32 values or 72 B of FP64 values and no column, one
    # 4-byte offset per slot position of a slice; slice pointers; x and y:
    # 4 or 8 B per DOF
    ptrs = 4 * ((n_dof // 3 + 63) // 64 + 1)
    n_sell = (by32 - 8.0 * n_dof - ptrs) / (36 + 4 / 64)
    assert abs(n_sell - round(n_sell)) < 1e-6 and round(n_sell) % 64 == 0 and n_sell >= n_dof // 3
    n_sell = round(n_sell)
    assert by64 == pytest.approx(n_sell * 72 + (n_sell // 64) * 4 + ptrs + 16.0 * n_dof, rel=1e-12)
