"""Captured iteration graphs across reassembly and buffer growth.

A context keeps its captured PCG iteration graphs while every address and
size they were captured with is unchanged (a reassembly writes new values
into the same buffers), and re-captures them when a buffer moves (a larger
mesh on the same context).  Each solve here must equal, bit for bit, the
same solve on a fresh context, with graphs and without (BHIP_NO_GRAPH=1 is
read per solve)."""
import os

import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem

pytestmark = pytest.mark.gpu

PRECS = [capi.PREC_FP64, capi.PREC_FP32_IR, capi.PREC_FP64_BLOCKED]


def _solve(ctx, prec):
    ctx.solve(tol=1e-10, precision=prec, maxit=20000)
    return ctx.solution()


def _fresh(prob, prec):
    with capi.Context(0) as c:
        setup_problem(c, prob)
        c.assemble()
        c.set_load(prob.load)
        c.build_rhs()
        return _solve(c, prec)


def _pair(kind):
    if kind == "hex":
        a, b = meshgen.elastic_cube(6), meshgen.elastic_cube(6)
        b.mat = b.mat.copy()
        b.mat[:, 0] *= 3.0  # other stiffness, same mesh and buffers
    else:
        a, b = meshgen.plane_stress_quad(16), meshgen.plane_stress_quad(16)
        b.mat = b.mat.copy()
        b.mat[:, [0, 1, 3]] *= 3.0
    return a, b


@pytest.mark.parametrize("kind", ["hex", "quad"])
@pytest.mark.parametrize("prec", PRECS)
def test_reassembly_reuses_graphs_bitwise(prec, kind):
    """The second and third assemblies write the SELL block values of the
    current layout themselves (no CSR-to-SELL pass): the same bits as a
    fresh context, which converts."""
    a, b = _pair(kind)
    print("reassembly", prec, kind, flush=True)
    with capi.Context(0) as ctx:
        setup_problem(ctx, a)
        ctx.assemble()
        ctx.set_load(a.load)
        ctx.build_rhs()
        xa = _solve(ctx, prec)
        ctx.set_materials(b.mat_kind, b.mat)
        ctx.assemble()
        ctx.build_rhs()
        xb = _solve(ctx, prec)
        # back to the first material: the same bits as the first solve
        ctx.set_materials(a.mat_kind, a.mat)
        ctx.assemble()
        ctx.build_rhs()
        xa2 = _solve(ctx, prec)
    assert np.array_equal(xa, _fresh(a, prec))
    assert np.array_equal(xb, _fresh(b, prec))
    assert np.array_equal(xa, xa2)
    assert not np.array_equal(xa, xb)


@pytest.mark.parametrize("prec", PRECS)
def test_growing_mesh_recaptures(prec):
    """Meshes of growing size on one context: every buffer the graphs use is
    reallocated at least once; each solve equals the fresh-context solve."""
    print("growing mesh", prec, flush=True)
    with capi.Context(0) as ctx:
        for n in (4, 7, 5, 9):
            p = meshgen.elastic_cube(n)
            setup_problem(ctx, p)
            ctx.assemble()
            ctx.set_load(p.load)
            ctx.build_rhs()
            x = _solve(ctx, prec)
            assert np.array_equal(x, _fresh(p, prec)), n


def test_graphs_and_eager_agree(monkeypatch):
    p = meshgen.elastic_cube(6)
    with capi.Context(0) as ctx:
        setup_problem(ctx, p)
        ctx.assemble()
        ctx.set_load(p.load)
        ctx.build_rhs()
        xg = _solve(ctx, capi.PREC_FP32_IR)
        monkeypatch.setenv("BHIP_NO_GRAPH", "1")
        ctx.assemble()
        ctx.build_rhs()
        xe = _solve(ctx, capi.PREC_FP32_IR)
        monkeypatch.delenv("BHIP_NO_GRAPH")
    assert os.environ.get("BHIP_NO_GRAPH") is None
    assert np.array_equal(xg, xe)
