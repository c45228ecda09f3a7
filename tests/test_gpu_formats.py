"""Matrix exchange through the C ABI on the GPU: the assembled matrix written
in the reference's formats, and systems the reference dumped loaded back and
solved on the device.

Checks and tolerances:
  - bhip_write_matrix (CSR ASCII, PARDISO upper 1-based) read by the
    reference's own readCSRMatrixInASCII equals the transpose of the oracle's
    lower-triangle matrix: pattern exact, values <= 1e-12 row-scaled
  - CSR binary written by the reference's writeCSRMatrixInBinary from the
    oracle's matrix, plus LoadVector.txt, loaded with bhip_read_matrix /
    bhip_read_vector: FP64 solve to 1e-12 within 1e-8 of the oracle's PCG;
    FP32-inner + refinement to 1e-10 within 1e-7
  - binary round trip is bit-exact; RCV text round trip (16 significant
    digits) within 1e-15 relative per entry
  - a loaded matrix has no mesh: assembly and loads are refused
"""
import numpy as np
import pytest

from betanoox_amd import capi, meshgen, setup_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _upper_of_lower(n, rp, ci, v):
    """Transpose a lower-triangle CSR (0-based) into the upper triangle."""
    import scipy.sparse as sp
    L = sp.csr_matrix((v, ci, rp), shape=(n, n))
    U = L.T.tocsr()
    U.sort_indices()
    return U.indptr.astype(np.int32), U.indices.astype(np.int32), U.data


def _need_ref():
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (make -C oracle ref needs the reference sources)")


def test_written_matrix_read_by_reference(hip_ctx_factory, tmp_path):
    _need_ref()
    prob = meshgen.elastic_cube(4)
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.assemble()
    path = tmp_path / "CSR_ASCII.txt"
    ctx.write_matrix(path, capi.FMT_CSR_ASCII)
    ia, ja, a = O.ref_read_csr(path, False)
    eq = O.assemble_problem(prob)
    rp, ci = eq.pattern()
    uia, uja, ua = _upper_of_lower(prob.n_dof, rp, ci, eq.values())
    assert np.array_equal(ia, uia + 1) and np.array_equal(ja, uja + 1)
    err = 0.0
    for i in range(prob.n_dof):
        s = slice(uia[i], uia[i + 1])
        err = max(err, np.abs(a[s] - ua[s]).max() / max(np.abs(ua[s]).max(), 1e-300))
    assert err <= 1e-12


def test_reference_dump_solved_on_device(hip_ctx_factory, tmp_path):
    _need_ref()
    prob = meshgen.elastic_cube(5)
    eq = O.assemble_problem(prob)
    rp, ci = eq.pattern()
    uia, uja, ua = _upper_of_lower(prob.n_dof, rp, ci, eq.values())
    kpath, bpath = tmp_path / "CSR_Binary.dat", tmp_path / "LoadVector.txt"
    O.ref_write_csr(kpath, True, uia + 1, uja + 1, ua)
    b = eq.rhs()
    capi.write_vector(bpath, capi.VEC_LOAD, b)
    st, x_o, it, _ = eq.pcg(b, eps=1e-13, maxit=20000, stop_rule=1)
    ctx = hip_ctx_factory()
    ctx.read_matrix(kpath, capi.FMT_CSR_BINARY)
    assert ctx.n_dof == prob.n_dof
    ctx.set_rhs(capi.read_vector(bpath, capi.VEC_LOAD, prob.n_dof))
    s = ctx.solve(tol=1e-12, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP64, maxit=20000)
    assert s["rel_resid"] <= 1e-12
    assert np.linalg.norm(ctx.solution() - x_o) <= 1e-8 * np.linalg.norm(x_o)
    s = ctx.solve(tol=1e-10, stop_rule=capi.STOP_REL_RESID, precision=capi.PREC_FP32_IR, maxit=20000)
    assert s["rel_resid"] <= 1e-10
    assert np.linalg.norm(ctx.solution() - x_o) <= 1e-7 * np.linalg.norm(x_o)
    # the lower export of the loaded matrix is the oracle's lower matrix
    rp2, ci2 = ctx.pattern_lower()
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci)
    assert np.array_equal(ctx.values_lower(), eq.values())


def test_round_trips_and_no_mesh(hip_ctx_factory, tmp_path):
    prob = meshgen.plane_stress_quad(8)
    ctx = setup_problem(hip_ctx_factory(), prob)
    ctx.assemble()
    v0 = ctx.values_lower()
    rp0, ci0 = ctx.pattern_lower()
    for fmt, rfmt, exact in [(capi.FMT_CSR_BINARY, capi.FMT_CSR_BINARY, True),
                             (capi.FMT_RCV_LOWER, capi.FMT_RCV, False),
                             (capi.FMT_RCV_UPPER, capi.FMT_RCV, False),
                             (capi.FMT_CSR_ASCII, capi.FMT_CSR_ASCII, False)]:
        path = tmp_path / ("m%d" % fmt)
        ctx.write_matrix(path, fmt)
        c2 = hip_ctx_factory()
        c2.read_matrix(path, rfmt)
        rp, ci = c2.pattern_lower()
        assert np.array_equal(rp, rp0) and np.array_equal(ci, ci0)
        v = c2.values_lower()
        if exact:
            assert np.array_equal(v, v0)
        else:
            assert np.all(np.abs(v - v0) <= 1e-15 * np.abs(v0))
    with pytest.raises(capi.BhipError) as e:
        c2.assemble()
    assert e.value.status == capi.ERR_STATE
    with pytest.raises(capi.BhipError) as e:
        c2.set_load(np.zeros(c2.n_dof))
    assert e.value.status == capi.ERR_STATE
    with pytest.raises(capi.BhipError) as e:
        c2.read_matrix(tmp_path / "missing", capi.FMT_CSR_BINARY)
    assert e.value.status == capi.ERR_ARG
