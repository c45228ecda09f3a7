"""Edge cases of the device path against the oracle: ragged DOF counts
(a node with no DOFs inside the numbering), every DOF restrained, a zero
right-hand side, a one-element mesh and a mesh without elements.

Tolerances: FP64 PCG solutions within 1e-8 of the oracle's FP64 PCG (both to
1e-12); FP32-inner refined solutions within 1e-8 (solved to 1e-11); exact
answers (restrained values, zeros) bit for bit on the FP64 CSR form.
"""
import types

import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PRECS = [capi.PREC_FP64, capi.PREC_FP32_IR, capi.PREC_FP64_BLOCKED]


def _solve(ctx, prec, tol):
    s = ctx.solve(tol=tol, stop_rule=capi.STOP_REL_RESID, precision=prec, inner_tol=1e-3, maxit=20000)
    return s, ctx.solution()


def _oracle_solution(prob):
    eq = O.assemble_problem(prob)
    st, x, _, _ = eq.pcg(eq.rhs(), eps=1e-12, stop_rule=1)
    assert st == 0
    return x


def _ragged(prob, at):
    """The same model with a node of no DOFs inserted at index `at`: the DOF
    numbering of every other node is unchanged (node.cpp:206-214 counts
    DOFs in node order), so restraints, prescribed values and loads are the
    same arrays."""
    xyz = np.insert(prob.xyz, at, [9.0, 9.0, 9.0], axis=0)
    conn = prob.conn + (prob.conn >= at)
    nd = np.full(xyz.shape[0], 3, np.int32)
    nd[at] = 0
    return xyz, conn.astype(np.int32), nd


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_FP32_IR])
def test_ragged_dof_counts(hip_ctx_factory, prec):
    """A node without DOFs in the middle of the numbering: same pattern, K and
    solution as the model without it, and against the oracle (whose DOF
    numbering takes the per-node counts).  The FP32 inner solve runs on the
    scalar CSR form (the 3x3-block form needs the same DOFs at every node),
    and the FP64 blocked precision is refused."""
    prob = meshgen.elastic_cube(4)
    print("case ragged", prec, flush=True)
    xyz, conn, nd = _ragged(prob, 37)
    ctx = hip_ctx_factory()
    assert ctx.set_mesh(3, xyz, conn, ndof_per_node=nd) == prob.n_dof
    ctx.set_restraints(prob.restrained, prob.prescribed)
    ctx.build_pattern()
    ctx.set_materials(prob.mat_kind, prob.mat, prob.extra.get("group"))
    ctx.set_integration(prob.order, prob.factor)
    ctx.assemble()
    ctx.set_load(prob.load)
    ctx.build_rhs()
    ref = setup_problem(hip_ctx_factory(), prob)
    ref.assemble()
    ref.set_load(prob.load)
    ref.build_rhs()
    rp, ci = ctx.pattern_lower()
    rp0, ci0 = ref.pattern_lower()
    assert np.array_equal(rp, rp0) and np.array_equal(ci, ci0)
    tol = 1e-12 if prec != capi.PREC_FP32_IR else 1e-11
    s, x = _solve(ctx, prec, tol)
    assert s["rel_resid"] <= tol
    x_o = _oracle_solution(prob)
    assert np.linalg.norm(x - x_o) <= 1e-8 * np.linalg.norm(x_o)
    _, x0 = _solve(ref, prec, tol)
    if prec == capi.PREC_FP64:  # the same CSR solve: the same bits
        assert np.array_equal(x, x0)
        with pytest.raises(capi.BhipError) as ei:
            ctx.solve(tol=1e-8, precision=capi.PREC_FP64_BLOCKED)
        assert ei.value.status == capi.ERR_ARG
    # the oracle's numbering with the per-node counts gives the same K
    oprob = types.SimpleNamespace(n_nodes=xyz.shape[0], ndof=nd, n_dof=prob.n_dof, restrained=prob.restrained,
                                  prescribed=prob.prescribed, conn=conn, xyz=xyz, mat_kind=prob.mat_kind,
                                  mat=prob.mat, order=prob.order, factor=prob.factor, load=prob.load,
                                  n_el=prob.n_el, extra=prob.extra)
    eq = O.assemble_problem(oprob, elem_group=prob.extra.get("group"))
    st, x_r, _, _ = eq.pcg(eq.rhs(), eps=1e-12, stop_rule=1)
    assert st == 0
    assert np.linalg.norm(x_r - x_o) <= 1e-14 * np.linalg.norm(x_o)


@pytest.mark.parametrize("prec", PRECS)
def test_every_dof_restrained(hip_ctx_factory, prec):
    """K is the identity (equation.cpp:500) and b the prescribed values: the
    solution is the prescribed field, bit for bit on the FP64 CSR form as in
    the oracle (alpha = r.z / p.q is exactly 1 there) and to 1e-12 relative
    on the blocked and refined forms (whose dot products are summed in
    another order)."""
    prob = meshgen.elastic_cube(3)
    print("case all restrained", prec, flush=True)
    rng = np.random.default_rng(11)
    prob.restrained = np.ones(prob.n_dof, np.uint8)
    prob.prescribed = rng.uniform(-1, 1, prob.n_dof)
    prob.load = np.zeros(prob.n_dof)
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.assemble()
    ctx.set_load(None)
    ctx.build_rhs()
    s, x = _solve(ctx, prec, 1e-12)
    if prec == capi.PREC_FP64:
        assert np.array_equal(x, prob.prescribed)
        assert np.array_equal(x, _oracle_solution(prob))
    else:
        assert np.linalg.norm(x - prob.prescribed) <= 1e-12 * np.linalg.norm(prob.prescribed)


@pytest.mark.parametrize("prec", PRECS)
def test_zero_right_hand_side(hip_ctx_factory, prec):
    """No load and zero prescribed values: x = 0 at once, no iteration."""
    prob = meshgen.elastic_cube(3)
    print("case zero rhs", prec, flush=True)
    prob.prescribed = np.zeros(prob.n_dof)
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.assemble()
    ctx.set_load(None)
    ctx.build_rhs()
    s, x = _solve(ctx, prec, 1e-10)
    assert s["iters"] == 0
    assert not np.any(x)


@pytest.mark.parametrize("prec", PRECS)
def test_one_element(hip_ctx_factory, prec):
    """A single hex8 (24 DOFs, 12 restrained) against the oracle."""
    prob = meshgen.elastic_cube(1)
    print("case one element", prec, flush=True)
    assert prob.n_el == 1
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.assemble()
    ctx.set_load(prob.load)
    ctx.build_rhs()
    tol = 1e-12 if prec != capi.PREC_FP32_IR else 1e-11
    s, x = _solve(ctx, prec, tol)
    x_o = _oracle_solution(prob)
    assert np.linalg.norm(x - x_o) <= 1e-8 * np.linalg.norm(x_o)


def test_mesh_without_elements(hip_ctx_factory):
    """Nodes and no elements: every row holds only its (zero) diagonal, so the
    solve reports the zero diagonal, and with zero-diagonal replacement
    (largeMatrix.cpp:43-79) K is the replacement value times I."""
    print("case no elements", flush=True)
    xyz = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    ctx = hip_ctx_factory()
    assert ctx.set_mesh(3, xyz, np.zeros((0, 8), np.int32)) == 6
    ctx.set_restraints(np.zeros(6, np.uint8))
    ctx.build_pattern()
    rp, ci = ctx.pattern_lower()
    assert np.array_equal(rp, np.arange(7)) and np.array_equal(ci, np.arange(6))
    ctx.set_materials(capi.MAT_ISO_ELASTIC, np.array([[1.0, 0.3]]))
    ctx.set_integration(2)
    ctx.assemble()
    f = np.arange(1.0, 7.0)
    ctx.set_load(f)
    ctx.build_rhs()
    with pytest.raises(capi.BhipError) as ei:
        ctx.solve(tol=1e-10)
    assert ei.value.status == capi.ERR_ZERO_DIAG
    ctx.set_zero_diag_replacement(2.0)
    s, x = _solve(ctx, capi.PREC_FP64, 1e-12)
    assert np.allclose(x, f / 2.0, rtol=1e-15, atol=0)
