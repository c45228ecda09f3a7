"""bhip_pcg_solve_noalloc (no reference counterpart; for callers that cannot
allocate device memory) against bhip_pcg_solve, called through the C ABI:
the same iterates bit for bit on two calls in a row, the second on a smaller
system, with and without an initial guess; no device memory is taken by the
call; zero diagonals are refused or replaced as by bhip_pcg_solve; other
precisions are refused."""
import ctypes as C

import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem

pytestmark = pytest.mark.gpu


def _assembled(n):
    prob = meshgen.elastic_cube(n)
    ctx = capi.Context(0)
    setup_problem(ctx, prob)
    ctx.assemble()
    ctx.set_load(prob.load)
    ctx.build_rhs()
    return ctx, ctx.rhs()


def _call(lib_fn, ctx, b, x0, opts):
    bb = np.ascontiguousarray(b, dtype=np.float64).copy()
    st = capi.Stats()
    x0p = None if x0 is None else capi._d(np.ascontiguousarray(x0, dtype=np.float64))
    rc = lib_fn(ctx.h, capi._d(bb), x0p, C.byref(opts), C.byref(st))
    return rc, bb, st.as_dict()


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_noalloc_matches_pcg_solve_twice_smaller_second():
    lib = capi.load_library()
    opts = capi.Context.opts(tol=1e-10, precision=capi.PREC_FP64, stop_rule=capi.STOP_REL_RESID, maxit=20000)
    for n, with_x0 in ((7, False), (4, True)):
        print("noalloc case n=%d x0=%s" % (n, with_x0), flush=True)
        ctx, b = _assembled(n)
        try:
            x0 = None
            if with_x0:
                x0 = np.random.default_rng(n).standard_normal(b.size) * 1e-3
            rc0, x_ref, s_ref = _call(lib.bhip_pcg_solve, ctx, b, x0, opts)
            assert rc0 == 0
            free0 = _free_bytes()
            rc1, x_na, s_na = _call(lib.bhip_pcg_solve_noalloc, ctx, b, x0, opts)
            free1 = _free_bytes()
            assert rc1 == 0, ctx.lib.bhip_last_error(ctx.h)
            assert free1 == free0
            assert np.array_equal(x_na, x_ref)
            assert s_na["iters"] == s_ref["iters"]
            assert s_na["rel_resid"] == s_ref["rel_resid"] and s_na["rel_resid"] <= 1e-10
        finally:
            ctx.close()


def test_noalloc_refusals():
    lib = capi.load_library()
    ctx, b = _assembled(3)
    try:
        o = capi.Context.opts(tol=1e-8, precision=capi.PREC_FP32_IR)
        rc, _, _ = _call(lib.bhip_pcg_solve_noalloc, ctx, b, None, o)
        assert rc == capi.ERR_UNSUPPORTED
        o = capi.Context.opts(tol=0.0, precision=capi.PREC_FP64)
        rc, _, _ = _call(lib.bhip_pcg_solve_noalloc, ctx, b, None, o)
        assert rc == capi.ERR_ARG
    finally:
        ctx.close()
    with capi.Context(0) as c2:  # nothing assembled
        o = capi.Context.opts(tol=1e-8, precision=capi.PREC_FP64)
        bb = np.zeros(3)
        st = capi.Stats()
        assert lib.bhip_pcg_solve_noalloc(c2.h, capi._d(bb), None, C.byref(o), C.byref(st)) == capi.ERR_STATE


def _with_isolated_node():
    """The 3^3 cube and one node outside every element: its three rows hold
    only a zero diagonal."""
    prob = meshgen.elastic_cube(3)
    xyz = np.vstack([prob.xyz, [[2.0, 2.0, 2.0]]])
    ctx = capi.Context(0)
    ctx.set_mesh(3, xyz, prob.conn)
    r = np.concatenate([prob.restrained, [0, 0, 0]])
    ctx.set_restraints(r, np.concatenate([prob.prescribed, [0.0, 0.0, 0.0]]))
    ctx.build_pattern()
    ctx.set_materials(prob.mat_kind, prob.mat, prob.extra.get("group"))
    ctx.set_integration(prob.order, prob.factor)
    ctx.assemble()
    ctx.set_load(np.concatenate([prob.load, [0.0, 0.0, 0.0]]))
    ctx.build_rhs()
    return ctx, ctx.rhs()


def test_noalloc_zero_diagonal():
    """Refused without a replacement value; with one, the same bits as
    bhip_pcg_solve on a second context (each replaces in its own values)."""
    print("noalloc case zero diagonal", flush=True)
    lib = capi.load_library()
    opts = capi.Context.opts(tol=1e-10, precision=capi.PREC_FP64, stop_rule=capi.STOP_REL_RESID, maxit=20000)
    ctx, b = _with_isolated_node()
    ref, b2 = _with_isolated_node()
    try:
        rc, _, _ = _call(lib.bhip_pcg_solve_noalloc, ctx, b, None, opts)
        assert rc == capi.ERR_ZERO_DIAG
        ctx.set_zero_diag_replacement(1.0)
        ref.set_zero_diag_replacement(1.0)
        free0 = _free_bytes()
        rc1, x_na, s_na = _call(lib.bhip_pcg_solve_noalloc, ctx, b, None, opts)
        assert _free_bytes() == free0
        assert rc1 == 0, ctx.lib.bhip_last_error(ctx.h)
        rc0, x_ref, s_ref = _call(lib.bhip_pcg_solve, ref, b2, None, opts)
        assert rc0 == 0
        assert np.array_equal(x_na, x_ref)
        assert s_na["iters"] == s_ref["iters"]
        assert not np.any(x_na[-3:])
    finally:
        ctx.close()
        ref.close()
