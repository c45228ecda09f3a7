"""Served request loop (bhip_server_*): RVE requests (hex8 unit cube with
inclusions, six affine boundary load cases) solved by several resident
contexts at once on one GPU.

Checks:
  - every served result is bitwise identical to bhip_solve_cases on a private
    context with the same options (same kernels, same reduction order), and
    within 1e-8 true relative residual
  - a request gives the same bits alone and in a mix of concurrent requests
  - requests packed into one batched solve (several packs of varying
    composition) give the same bits as their private solves
  - a slot reuses its pattern for a request of the same topology, with the
    same bits as a fresh build
  - admission: with a budget that holds one request at a time, no two
    requests overlap in time; a request above the budget is refused
  - a failing request (inverted element) reports its status and message and
    the server keeps serving
  - a server run in BHIP_PREC_FP64 or BHIP_PREC_FP64_BLOCKED solves each
    request's cases in turn with the same bits as the private solve
  - a served request (FP32-inner server refined to 1e-11) against the
    oracle's FP64 solve of each load case to 1e-13: relative error <= 1e-8
"""
import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-8


def _direct(hip_ctx_factory, prob):
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.assemble(capi.ASM_STIFFNESS)
    o = capi.Context.opts(tol=TOL, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR, inner_tol=1e-3,
                          maxit=40000, max_outer=0, check_every=0)
    X, st = ctx.solve_cases(np.stack(prob.extra["cases"]), opts=o)
    return X, st


def test_served_equals_direct_and_mix(hip_ctx_factory):
    probs = [meshgen.rve_request(n, 100 + n) for n in (6, 9, 12, 7, 10, 8)]
    print("requests", [p.n_dof for p in probs], flush=True)
    ref = [_direct(hip_ctx_factory, p) for p in probs]
    with capi.Server(0, max_slots=4, max_pack=1) as srv:
        # alone
        rid = srv.submit(probs[2])
        X_alone, s_alone = srv.wait(rid)
        assert s_alone["status"] == capi.OK
        assert np.array_equal(X_alone, ref[2][0])
        # in a mix, all submitted at once
        ids = [srv.submit(p) for p in probs]
        out = [srv.wait(i) for i in ids]
    for (X, s), (Xr, sr) in zip(out, ref):
        assert s["status"] == capi.OK and s["rel_resid"] <= TOL
        assert np.array_equal(X, Xr)
    assert np.array_equal(out[2][0], X_alone)
    # the mix ran on more than one slot and overlapped in time
    assert len({s["slot"] for _, s in out}) > 1
    spans = sorted((s["start_ms"], s["end_ms"]) for _, s in out)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))


def test_packed_requests_same_bits(hip_ctx_factory):
    """Requests queued together are solved in packs; whatever the packs'
    composition, every result equals the private solve bit for bit."""
    probs = [meshgen.rve_request(n, 600 + n) for n in (6, 9, 12, 7, 10, 8, 11)]
    ref = [_direct(hip_ctx_factory, p) for p in probs]
    with capi.Server(0, max_slots=1, max_pack=8) as srv:
        ids = [srv.submit(p) for p in probs]
        out = [srv.wait(i) for i in ids]
    sizes = [s["pack_size"] for _, s in out]
    print("pack sizes", sizes, flush=True)
    assert max(sizes) > 1
    for (X, s), (Xr, sr) in zip(out, ref):
        assert s["status"] == capi.OK and s["rel_resid"] <= TOL
        assert s["iters"] == sr["iters"]
        assert np.array_equal(X, Xr)
    with capi.Server(0, max_slots=3, max_pack=4) as srv:
        ids = [srv.submit(p) for p in probs[::-1]]
        out2 = [srv.wait(i) for i in ids]
    for (X, s), (Xr, _) in zip(out2, ref[::-1]):
        assert s["status"] == capi.OK
        assert np.array_equal(X, Xr)


def test_pattern_reuse_same_bits(hip_ctx_factory):
    a = meshgen.rve_request(9, 1)
    b = meshgen.rve_request(9, 2)  # same topology, other inclusions
    with capi.Server(0, max_slots=1) as srv:
        Xa, sa = srv.wait(srv.submit(a))
        Xb, sb = srv.wait(srv.submit(b))
        Xa2, sa2 = srv.wait(srv.submit(a))
    assert sa["reused_pattern"] == 0 and sb["reused_pattern"] == 1 and sa2["reused_pattern"] == 1
    assert np.array_equal(Xa, Xa2)
    assert np.array_equal(Xb, _direct(hip_ctx_factory, b)[0])


def test_admission_by_footprint():
    probs = [meshgen.rve_request(8, 200 + k) for k in range(4)]
    lib = capi.load_library()
    need = lib.bhip_request_footprint(probs[0].n_nodes, probs[0].n_el, 6)
    with capi.Server(0, max_slots=4, hbm_budget=1.5 * need) as srv:
        ids = [srv.submit(p) for p in probs]
        out = [srv.wait(i) for i in ids]
        big = meshgen.rve_request(24, 9)
        with pytest.raises(capi.BhipError) as e:
            srv.submit(big)
        assert e.value.status == capi.ERR_CAPACITY
    spans = sorted((s["start_ms"], s["end_ms"]) for _, s in out)
    for i in range(len(spans) - 1):
        assert spans[i + 1][0] >= spans[i][1]
    # served in arrival order
    starts = [s["start_ms"] for _, s in out]
    assert starts == sorted(starts)


def test_failing_request_reported(hip_ctx_factory):
    good = meshgen.rve_request(6, 4)
    bad = meshgen.rve_request(6, 5)
    bad.conn = bad.conn.copy()
    bad.conn[0] = bad.conn[0][[4, 5, 6, 7, 0, 1, 2, 3]]  # mirrored element: negative Jacobian
    with capi.Server(0, max_slots=2) as srv:
        ib = srv.submit(bad)
        ig = srv.submit(good)
        with pytest.raises(capi.BhipError) as e:
            srv.wait(ib)
        assert e.value.status == capi.ERR_NEG_DETJ
        X, s = srv.wait(ig)
        assert s["status"] == capi.OK
        X2, _ = srv.wait(srv.submit(good))
    assert np.array_equal(X, _direct(hip_ctx_factory, good)[0])
    assert np.array_equal(X2, X)


def test_other_case_counts(hip_ctx_factory):
    """Requests with 3 cases pack together; a request with 10 cases (more than
    one case group) is solved alone in groups of 8 + 2; both equal their
    private solves bit for bit."""
    small = [meshgen.rve_request(n, 700 + n) for n in (6, 8, 7)]
    cases3 = [np.stack(p.extra["cases"])[:3] for p in small]
    big = meshgen.rve_request(7, 777)
    c6 = np.stack(big.extra["cases"])
    cases10 = np.concatenate([c6, 0.5 * c6[:4]])
    o = capi.Context.opts(tol=TOL, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR, inner_tol=1e-3,
                          maxit=40000, max_outer=0, check_every=0)
    ref3 = []
    for p, c in zip(small, cases3):
        ctx = setup_problem(hip_ctx_factory(), p)
        ctx.assemble(capi.ASM_STIFFNESS)
        ref3.append(ctx.solve_cases(c, opts=o)[0])
    ctx = setup_problem(hip_ctx_factory(), big)
    ctx.assemble(capi.ASM_STIFFNESS)
    ref10 = np.concatenate([ctx.solve_cases(cases10[:8], opts=o)[0], ctx.solve_cases(cases10[8:], opts=o)[0]])
    with capi.Server(0, max_slots=1, max_pack=8) as srv:
        ib = srv.submit(big, cases=cases10)
        ids = [srv.submit(p, cases=c) for p, c in zip(small, cases3)]
        Xb, sb = srv.wait(ib)
        out = [srv.wait(i) for i in ids]
    assert sb["status"] == capi.OK and sb["pack_size"] == 1 and sb["rel_resid"] <= TOL
    assert np.array_equal(Xb, ref10)
    for (X, s), Xr in zip(out, ref3):
        assert s["status"] == capi.OK and s["rel_resid"] <= TOL
        assert np.array_equal(X, Xr)
    assert max(s["pack_size"] for _, s in out) > 1


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_FP64_BLOCKED])
def test_fp64_precisions_served(hip_ctx_factory, prec):
    probs = [meshgen.rve_request(n, 900 + n) for n in (6, 8)]
    o = capi.Context.opts(tol=TOL, stop_rule=capi.STOP_REL_RESID, precision=prec, inner_tol=1e-3,
                          maxit=40000, max_outer=0, check_every=0)
    ref = []
    for p in probs:
        ctx = setup_problem(hip_ctx_factory(), p)
        ctx.assemble(capi.ASM_STIFFNESS)
        ref.append(ctx.solve_cases(np.stack(p.extra["cases"]), opts=o)[0])
    with capi.Server(0, max_slots=2, precision=prec) as srv:
        ids = [srv.submit(p) for p in probs]
        out = [srv.wait(i) for i in ids]
    for (X, st), Xr in zip(out, ref):
        assert st["status"] == capi.OK and st["rel_resid"] <= TOL and st["pack_size"] == 1
        assert np.array_equal(X, Xr)


def test_served_request_against_oracle():
    """Served displacements of an RVE request (served alone and beside three
    other requests, FP32-inner refined to 1e-11) against the oracle: the
    reference's assembly and FP64 PCG to 1e-13 for each load case."""
    probs = [meshgen.rve_request(n, 300 + n) for n in (7, 5, 9, 6)]
    with capi.Server(0, max_slots=4, tol=1e-11) as srv:
        ids = [srv.submit(p) for p in probs]
        out = [srv.wait(i) for i in ids]
    for prob, (X, st) in zip(probs[:2], out[:2]):
        assert st["status"] == capi.OK and st["rel_resid"] <= 1e-11
        P = np.stack(prob.extra["cases"])
        for k in range(P.shape[0]):
            prob.prescribed = P[k]
            eq = O.assemble_problem(prob, elem_group=prob.extra["group"])
            status, x_o, _, _ = eq.pcg(eq.rhs(), eps=1e-13, maxit=20000, stop_rule=1)
            assert status == 0
            assert np.linalg.norm(X[k] - x_o) <= 1e-8 * np.linalg.norm(x_o), k
