"""Packed batched solve (bhip_pack_solve_cases) through the C ABI: the load
cases of several assembled RVE systems (hex8, six affine boundary load cases)
advance together in one persistent set of kernels.

Checks:
  - every system's solution in a pack is bitwise identical to
    bhip_solve_cases on that context alone (same records, same order), and
    within 1e-8 true relative residual
  - the pack order does not change any system's bits
  - a pack that mixes small systems (four waves per slice) with a large one
    (one wave per slice, 2,327 slices) and an unequal number of refinement
    steps per system gives the same bits as the systems alone
  - nrhs other than six; a context that cannot join a pack (plane stress,
    2 DOFs per node) is reported as unsupported while the others are solved
"""
import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem

pytestmark = pytest.mark.gpu

TOL = 1e-8


def _opts():
    return capi.Context.opts(tol=TOL, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR, inner_tol=1e-3,
                             maxit=40000, max_outer=0, check_every=0)


def _assembled(factory, prob):
    ctx = setup_problem(factory(), prob)
    ctx.assemble(capi.ASM_STIFFNESS)
    return ctx


def _alone(factory, prob, cases):
    ctx = _assembled(factory, prob)
    X, st = ctx.solve_cases(cases, opts=_opts())
    return X, st


def test_pack_equals_alone_and_order(hip_ctx_factory):
    probs = [meshgen.rve_request(n, 300 + n) for n in (6, 11, 16, 9, 20)]
    cases = [np.stack(p.extra["cases"]) for p in probs]
    print("systems", [p.n_dof for p in probs], flush=True)
    ref = [_alone(hip_ctx_factory, p, c) for p, c in zip(probs, cases)]
    ctxs = [_assembled(hip_ctx_factory, p) for p in probs]
    with capi.Pack(0) as pk:
        X, S, status = pk.solve_cases(ctxs, cases, opts=_opts())
        assert status == [capi.OK] * len(probs)
        for (Xr, sr), x, s in zip(ref, X, S):
            assert s["rel_resid"] <= TOL
            assert s["iters"] == sr["iters"] and s["outer_iters"] == sr["outer_iters"]
            assert np.array_equal(x, Xr)
        perm = [3, 0, 4, 2, 1]
        X2, S2, status2 = pk.solve_cases([ctxs[i] for i in perm], [cases[i] for i in perm], opts=_opts())
        assert status2 == [capi.OK] * len(probs)
        for j, i in enumerate(perm):
            assert np.array_equal(X2[j], ref[i][0])


def test_pack_mixed_split_modes(hip_ctx_factory):
    small = meshgen.rve_request(8, 41)
    large = meshgen.rve_request(52, 42)  # 148,877 nodes: 2,327 slices, one wave per slice
    probs = [small, large, meshgen.rve_request(12, 43)]
    cases = [np.stack(p.extra["cases"]) for p in probs]
    print("systems", [p.n_dof for p in probs], flush=True)
    ref = [_alone(hip_ctx_factory, p, c) for p, c in zip(probs, cases)]
    ctxs = [_assembled(hip_ctx_factory, p) for p in probs]
    with capi.Pack(0) as pk:
        X, S, status = pk.solve_cases(ctxs, cases, opts=_opts())
    assert status == [capi.OK] * 3
    for (Xr, sr), x, s in zip(ref, X, S):
        assert s["rel_resid"] <= TOL
        assert np.array_equal(x, Xr)


def test_pack_other_nrhs_and_unsupported(hip_ctx_factory):
    probs = [meshgen.rve_request(n, 500 + n) for n in (7, 10)]
    cases = [np.stack(p.extra["cases"])[:3] for p in probs]
    ref = [_alone(hip_ctx_factory, p, c) for p, c in zip(probs, cases)]
    quad = meshgen.plane_stress_quad(8)
    qctx = setup_problem(hip_ctx_factory(), quad)
    qctx.assemble(capi.ASM_STIFFNESS)
    qcases = np.stack([quad.prescribed] * 3)
    ctxs = [_assembled(hip_ctx_factory, p) for p in probs]
    with capi.Pack(0) as pk:
        X, S, status = pk.solve_cases([ctxs[0], qctx, ctxs[1]], [cases[0], qcases, cases[1]], opts=_opts())
    assert status[1] == capi.ERR_UNSUPPORTED
    assert status[0] == capi.OK and status[2] == capi.OK
    assert np.array_equal(X[0], ref[0][0])
    assert np.array_equal(X[2], ref[1][0])
