"""Resident transient diffusion (bhip_heat_run) against the oracle and a
closed form.

Checks and tolerances:
  - steps against the oracle, which reassembles K and M and the lift every
    step as the reference does (DiffusionModel.cpp:711-831) and solves each
    step with FP64 PCG to 1e-13: nonzero boundary increments and a nonzero
    boundary temperature exercise both lift terms; FP64 steps at 1e-12 agree
    to 1e-9 relative, FP32-inner steps at 1e-10 to 1e-7
  - separable sine mode, uniform k and capacity: the mode is an exact
    eigenvector of the hex8 K and consistent M (2-point Gauss is exact), so
    backward Euler scales it by 1/(1 + 3 mu dt) per step with
    mu = 6 (1 - cos(pi h)) / (h^2 (2 + cos(pi h))); agreement to 1e-9
  - the four initial-guess modes give the same temperatures (1e-8); the
    projected guess takes fewer iterations than the previous increment
  - a run without a heat-transient assembly is refused
"""
import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bumped(n, seed=None):
    """Diffusion cube whose boundary carries a nonzero temperature and a
    nonzero prescribed increment every step."""
    prob = meshgen.diffusion_cube(n, seed=seed)
    u0 = prob.extra["u0"].copy()
    bnd = prob.restrained.astype(bool)
    u0[bnd] = 0.2 + 0.1 * prob.xyz[bnd, 0]
    prob.prescribed = np.where(bnd, 0.01 * (1.0 + prob.xyz[:, 1]), 0.0)
    return prob, u0


def _oracle_steps(prob, u0, alpha, dt, steps):
    u = u0.copy()
    for _ in range(steps):
        eq = O.assemble_heat(prob, alpha, dt, u)
        st, du, it, nrm = eq.pcg(eq.rhs(), eps=1e-13, maxit=20000, stop_rule=1)
        assert st == 0
        u = u + du
    return u


@pytest.mark.parametrize("alpha", [1.0, 0.5])
def test_heat_steps_vs_oracle(hip_ctx_factory, alpha):
    prob, u0 = _bumped(7)
    dt, steps = 2e-3, 4
    print("case heat7 alpha", alpha, flush=True)
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.set_restraints(prob.restrained, prob.prescribed)
    ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[alpha, dt], u_state=u0)
    u_ref = _oracle_steps(prob, u0, alpha, dt, steps)
    u64, s64 = ctx.heat_run(steps, u0, capi.HEAT_X0_REFERENCE, tol=1e-12, stop_rule=capi.STOP_REL_RESID,
                            precision=capi.PREC_FP64, maxit=20000)
    assert s64["steps"] == steps and s64["max_rel_resid"] <= 1e-12
    scale = np.abs(u_ref).max()
    assert np.abs(u64 - u_ref).max() <= 1e-9 * scale
    u32, s32 = ctx.heat_run(steps, u0, capi.HEAT_X0_PREVIOUS, tol=1e-10, stop_rule=capi.STOP_REL_RESID,
                            precision=capi.PREC_FP32_IR, maxit=20000)
    assert s32["max_rel_resid"] <= 1e-10
    assert np.abs(u32 - u_ref).max() <= 1e-7 * scale


def test_heat_sine_mode_decay(hip_ctx_factory):
    n, dt, steps = 10, 5e-3, 6
    prob = meshgen.diffusion_cube(n)
    prob.mat = np.array([[1.0, 1.0, 1.0, 1.0]])
    x, y, z = prob.xyz.T
    u0 = np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)
    u0[prob.restrained.astype(bool)] = 0.0
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[1.0, dt], u_state=u0)
    u, st = ctx.heat_run(steps, u0, capi.HEAT_X0_ZERO, tol=1e-12, stop_rule=capi.STOP_REL_RESID,
                         precision=capi.PREC_FP64, maxit=20000)
    h = 1.0 / n
    c = np.cos(np.pi * h)
    mu = 6.0 * (1.0 - c) / (h * h * (2.0 + c))
    exact = u0 / (1.0 + 3.0 * mu * dt) ** steps
    assert np.abs(u - exact).max() <= 1e-9 * np.abs(u0).max()


def test_heat_x0_modes_agree(hip_ctx_factory):
    """The four initial-guess modes reach the same field (1e-8), and the
    projected guess (whose basis of up to 16 increments restarts at step 17)
    needs fewer inner iterations than the previous increment over 20 steps."""
    prob, u0 = _bumped(6)
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.set_restraints(prob.restrained, prob.prescribed)
    ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[1.0, 1e-3], u_state=u0)
    outs, iters = [], {}
    for mode in (capi.HEAT_X0_ZERO, capi.HEAT_X0_PREVIOUS, capi.HEAT_X0_REFERENCE, capi.HEAT_X0_PROJECTED):
        u, st = ctx.heat_run(5, u0, mode, tol=1e-10, stop_rule=capi.STOP_REL_RESID,
                             precision=capi.PREC_FP32_IR, maxit=20000)
        outs.append(u)
    for u in outs[1:]:
        assert np.abs(u - outs[0]).max() <= 1e-8 * np.abs(outs[0]).max()
    for mode in (capi.HEAT_X0_PREVIOUS, capi.HEAT_X0_PROJECTED):
        u, st = ctx.heat_run(20, u0, mode, tol=1e-10, stop_rule=capi.STOP_REL_RESID,
                             precision=capi.PREC_FP32_IR, maxit=20000)
        iters[mode] = (st["iters"], u)
    assert np.abs(iters[capi.HEAT_X0_PROJECTED][1] - iters[capi.HEAT_X0_PREVIOUS][1]).max() <= \
        1e-8 * np.abs(iters[capi.HEAT_X0_PREVIOUS][1]).max()
    assert iters[capi.HEAT_X0_PROJECTED][0] < iters[capi.HEAT_X0_PREVIOUS][0]


def test_heat_projected_guess_fp64_and_csr(hip_ctx_factory):
    """The projected guess on the plain FP64 solve (Mbar du on the FP64 CSR)
    against the oracle's per-step reassembly, and repeated bit for bit."""
    prob, u0 = _bumped(5, seed=2)
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.set_restraints(prob.restrained, prob.prescribed)
    ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[1.0, 1e-3], u_state=u0)
    u1, s1 = ctx.heat_run(12, u0, capi.HEAT_X0_PROJECTED, tol=1e-12, stop_rule=capi.STOP_REL_RESID,
                          precision=capi.PREC_FP64, maxit=20000)
    u2, _ = ctx.heat_run(12, u0, capi.HEAT_X0_PROJECTED, tol=1e-12, stop_rule=capi.STOP_REL_RESID,
                         precision=capi.PREC_FP64, maxit=20000)
    assert np.array_equal(u1, u2)
    u3, s3 = ctx.heat_run(12, u0, capi.HEAT_X0_PREVIOUS, tol=1e-12, stop_rule=capi.STOP_REL_RESID,
                          precision=capi.PREC_FP64, maxit=20000)
    u_ref = _oracle_steps(prob, u0, 1.0, 1e-3, 12)
    assert np.abs(u1 - u_ref).max() <= 1e-9 * np.abs(u_ref).max()
    assert np.abs(u1 - u3).max() <= 1e-9 * np.abs(u3).max()
    assert s1["iters"] < s3["iters"]


def test_heat_run_needs_heat_assembly(hip_ctx_factory):
    prob = meshgen.elastic_cube(3)
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.assemble(capi.ASM_STIFFNESS)
    with pytest.raises(capi.BhipError) as e:
        ctx.heat_run(1, np.zeros(prob.n_dof))
    assert e.value.status == capi.ERR_STATE


def test_refined_heat_run_on_runs_and_csr(hip_ctx_factory):
    """The refined heat run builds -K u_s from an FP64 column-run copy of K
    when the layout is the run form, and from the CSR of K otherwise
    (BHIP_SELL_RUNS=0).  Both forms are exercised (the run layout moves
    fewer SpMV bytes), agree with the oracle's per-step reassembly (1e-8 at
    a 1e-11 solve) and with each other (1e-9)."""
    import os
    prob, u0 = _bumped(7, seed=3)
    dt, steps = 2e-3, 3
    u_ref = _oracle_steps(prob, u0, 1.0, dt, steps)
    res = {}
    for runs in (True, False):
        old = os.environ.get("BHIP_SELL_RUNS")
        os.environ["BHIP_SELL_RUNS"] = "1" if runs else "0"
        try:
            ctx = setup_problem(hip_ctx_factory(), prob)
            ctx.set_restraints(prob.restrained, prob.prescribed)
            ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[1.0, dt], u_state=u0)
            _, by = ctx.bench_spmv(fp32=True, reps=2)  # builds the layout under the switch
            u, st = ctx.heat_run(steps, u0, capi.HEAT_X0_PREVIOUS, tol=1e-11, stop_rule=capi.STOP_REL_RESID,
                                 precision=capi.PREC_FP32_IR, inner_tol=1e-3, maxit=20000)
        finally:
            if old is None:
                del os.environ["BHIP_SELL_RUNS"]
            else:
                os.environ["BHIP_SELL_RUNS"] = old
        assert st["steps"] == steps and st["max_rel_resid"] <= 1e-11
        res[runs] = (u, by)
    assert res[True][1] < res[False][1]
    scale = np.abs(u_ref).max()
    for runs in (True, False):
        assert np.abs(res[runs][0] - u_ref).max() <= 1e-8 * scale, runs
    assert np.abs(res[True][0] - res[False][0]).max() <= 1e-9 * scale
