"""Pins of the oracle (CPU, no GPU): each check is chosen so that a plausible
mistake in the restatement (dropped term, wrong sign or index, transposed
operand) fails it.  Sources of truth: the reference's own debug-main known
answers (matrix.cpp:687-696), the reference's own shape.cpp built into
oracle/_ref, closed forms, invariants and brute force on tiny inputs."""
import ctypes as C
import json
import os
import math

import numpy as np
import pytest

from betanoox_amd import capi, meshgen
from oracle import oracle as O


# ----------------------------------------------------------------- quadrature
def test_gauss_two_point_closed_form():
    c, w = O.gauss(2)
    assert np.allclose(c, [-1 / math.sqrt(3), 1 / math.sqrt(3)], rtol=0, atol=2e-16)
    assert np.allclose(w, [1.0, 1.0], rtol=0, atol=4e-16)


@pytest.mark.parametrize("k", [1, 3, 4, 5])
def test_gauss_matches_legendre(k):
    c, w = O.gauss(k)
    cr, wr = np.polynomial.legendre.leggauss(k)
    assert np.allclose(c, cr, atol=1e-14) and np.allclose(w, wr, atol=1e-14)


def test_hex_points_order_xi_outer_zeta_inner():
    xi, eta, zeta, w = O.hex_points(2)
    g = 1 / math.sqrt(3)
    # gauss.cpp:331-349: ip = (i*k + j)*k + l, xi outer, zeta inner
    assert np.allclose(xi, [-g, -g, -g, -g, g, g, g, g])
    assert np.allclose(eta, [-g, -g, g, g, -g, -g, g, g])
    assert np.allclose(zeta, [-g, g, -g, g, -g, g, -g, g])
    assert np.allclose(w, 1.0)


# ---------------------------------------------------------- shape functions
HEX_NODES = np.array([[-1, -1, -1], [-1, 1, -1], [-1, 1, 1], [-1, -1, 1],
                      [1, -1, -1], [1, 1, -1], [1, 1, 1], [1, -1, 1]], float)


def _shape(pt):
    S = np.zeros(8)
    O.lib().orc_shape_hex8(O.d(S), *pt)
    return S


def _deriv(pt):
    a, b, c = np.zeros(8), np.zeros(8), np.zeros(8)
    O.lib().orc_deriv_hex8(O.d(a), O.d(b), O.d(c), *pt)
    return np.stack([a, b, c])


def test_hex8_kronecker_and_partition_of_unity():
    for i, nd in enumerate(HEX_NODES):
        S = _shape(nd)
        assert np.allclose(S, np.eye(8)[i], atol=1e-15)
    rng = np.random.default_rng(1)
    for pt in rng.uniform(-1, 1, (20, 3)):
        assert abs(_shape(pt).sum() - 1) < 1e-15
        assert np.allclose(_deriv(pt).sum(axis=1), 0, atol=1e-15)


def test_hex8_derivative_is_gradient_of_reference_shape():
    """dN/dxi of the oracle equals the central difference of the reference's
    own shapeHex8 and, bit for bit, the reference's own derivHex8 (both from
    oracle/_ref, built from shape.cpp)."""
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    rng = np.random.default_rng(2)
    h = 1e-6
    for pt in rng.uniform(-1, 1, (10, 3)):
        D = _deriv(pt)
        for axis in range(3):
            p1, p0 = pt.copy(), pt.copy()
            p1[axis] += h
            p0[axis] -= h
            S1, S0 = np.zeros(8), np.zeros(8)
            R._Z9shapeHex8Pdddd(O.d(S1), *p1)
            R._Z9shapeHex8Pdddd(O.d(S0), *p0)
            assert np.allclose((S1 - S0) / (2 * h), D[axis], atol=1e-9)
        a, b, c = np.zeros(8), np.zeros(8), np.zeros(8)
        R._Z9derivHex8PdS_S_ddd(O.d(a), O.d(b), O.d(c), *pt)
        assert np.array_equal(np.stack([a, b, c]), D)
        S, Sr = _shape(pt), np.zeros(8)
        R._Z9shapeHex8Pdddd(O.d(Sr), *pt)
        assert np.array_equal(S, Sr)


def test_quad4_bit_exact_with_reference():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(3)
    for pt in rng.uniform(-1, 1, (10, 2)):
        S, Sr = np.zeros(4), np.zeros(4)
        O.lib().orc_shape_quad4(O.d(S), *pt)
        R._Z10shapeQuad4Pddd(O.d(Sr), *pt)
        assert np.array_equal(S, Sr)
        a, b, ar, br = np.zeros(4), np.zeros(4), np.zeros(4), np.zeros(4)
        O.lib().orc_deriv_quad4(O.d(a), O.d(b), *pt)
        R._Z10derivQuad4PdS_dd(O.d(ar), O.d(br), *pt)
        assert np.array_equal(a, ar) and np.array_equal(b, br)


# ------------------------------------------------------------------ dense LU
# Known answers of the reference's dense-matrix debug main (matrix.cpp:662-696),
# kept in tests/golden/matrix_debug_main.json (written by the script beside it).
GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                     "matrix_debug_main.json")))


def test_invert_matches_reference_debug_main():
    """matrix.cpp:696: B = {{2,4},{7,-3}}, B.Invert = {{0.088235,0.11765},{0.20588,-.058824}}
    (printed to 5 significant digits)."""
    st, Bi = O.invert(np.array(GOLDEN["B"]))
    assert st == 0
    assert np.allclose(Bi, GOLDEN["invert"], rtol=5e-5, atol=0)
    assert np.allclose(Bi @ np.array(GOLDEN["B"]), np.eye(2), atol=1e-15)


def test_ludcmp_matches_reference_debug_main():
    """matrix.cpp:692: A = {{1,2},{3,1}} factors in place to {{3,1},{.3333,1.6667}}
    (printed to 4 decimals) and A.Solve({5,5}) gives (1,2)."""
    A = np.array(GOLDEN["A"])
    indx = np.zeros(2, np.int32)
    dd = C.c_double()
    O.lib().orc_ludcmp(2, O.d(A), O.i32(indx), C.byref(dd))
    assert np.allclose(A, GOLDEN["lu_factors"], atol=5e-5)
    assert np.allclose(A, [[3, 1], [1 / 3, 5 / 3]], atol=1e-15)  # the exact factors
    b = np.array(GOLDEN["solve_rhs"])
    O.lib().orc_lubksb(2, O.d(A), O.i32(indx), O.d(b))
    assert np.allclose(b, GOLDEN["solve_x"], atol=1e-15)


def test_invert_random_against_numpy():
    rng = np.random.default_rng(4)
    for n in (3, 6):
        M = rng.standard_normal((n, n)) + n * np.eye(n)
        st, Mi = O.invert(M)
        assert np.allclose(Mi @ M, np.eye(n), atol=1e-13)


# ----------------------------------------------------------------- materials
@pytest.mark.parametrize("E,nu", [(1.0, 0.3), (70.0, 0.22), (3.0, 0.35), (42.0, 0.0)])
def test_iso_cmat_equals_lame(E, nu):
    Cm = O.iso_cmat(E, nu)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    ref = np.zeros((6, 6))
    ref[:3, :3] = lam
    ref[np.arange(3), np.arange(3)] = lam + 2 * mu
    ref[np.arange(3, 6), np.arange(3, 6)] = mu
    assert np.allclose(Cm, ref, rtol=2e-15, atol=2e-15 * E)


def test_plane_stress_closed_form():
    E, nu = 1.0, 0.3
    Cm = O.plane_stress_cmat(E, E, nu, E / (2 * (1 + nu)))
    ref = E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    assert np.allclose(Cm, ref, atol=1e-15)


# ------------------------------------------------------------ element matrices
def _affine_hex(rng, scale=1.0):
    A = np.eye(3) * scale + 0.2 * scale * rng.standard_normal((3, 3))
    t = rng.standard_normal(3)
    X = (HEX_NODES + 1) / 2 @ A.T + t
    return X, abs(np.linalg.det(A))


def test_hex8_rigid_modes_and_rank():
    rng = np.random.default_rng(5)
    X = (HEX_NODES + 1) / 2 * 0.3 + rng.uniform(-0.02, 0.02, (8, 3)) + 5.0
    st, K = O.ke_hex8(X, O.iso_cmat(2.0, 0.3))
    assert st == 0
    assert np.allclose(K, K.T, atol=0)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-12 * ev.max()) == 6 and ev.min() > -1e-12 * ev.max()
    for k in range(6):  # 3 translations + 3 rotations
        u = np.zeros((8, 3))
        if k < 3:
            u[:, k] = 1
        else:
            w = np.zeros(3)
            w[k - 3] = 1
            u = np.cross(w, X - X.mean(0))
        assert np.abs(K @ u.ravel()).max() < 1e-12 * np.abs(K).max()


def test_hex8_energy_of_uniform_strain_is_exact():
    """u = eps.x on an affine brick: u^T K u = V eps^T C eps (2x2x2 Gauss is
    exact here).  Catches wrong B rows, shear factors, C transposition."""
    rng = np.random.default_rng(6)
    for _ in range(5):
        X, V = _affine_hex(rng, scale=rng.uniform(0.1, 2.0))
        E, nu = rng.uniform(1, 100), rng.uniform(0.1, 0.4)
        Cm = O.iso_cmat(E, nu)
        st, K = O.ke_hex8(X, Cm)
        assert st == 0
        G = rng.standard_normal((3, 3))
        eps = 0.5 * (G + G.T)
        u = X @ eps.T
        ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[0, 2]])
        assert np.isclose(u.ravel() @ K @ u.ravel(), V * ev @ Cm @ ev, rtol=1e-12)


def test_quad4_energy_of_uniform_strain_is_exact():
    rng = np.random.default_rng(7)
    for _ in range(5):
        A = np.eye(2) + 0.2 * rng.standard_normal((2, 2))
        X = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float) @ A.T + rng.standard_normal(2)
        Cm = O.plane_stress_cmat(2.0, 2.0, 0.3, 2.0 / 2.6)
        t = rng.uniform(0.5, 2)
        st, K = O.ke_quad4(X, Cm, factor=t)
        G = rng.standard_normal((2, 2))
        eps = 0.5 * (G + G.T)
        u = X @ eps.T
        ev = np.array([eps[0, 0], eps[1, 1], 2 * eps[0, 1]])
        assert np.isclose(u.ravel() @ K @ u.ravel(), t * abs(np.linalg.det(A)) * ev @ Cm @ ev, rtol=1e-12)


def test_heat_element_invariants():
    rng = np.random.default_rng(8)
    X, V = _affine_hex(rng)
    k = np.array([2.0, 3.0, 5.0])
    st, K, M = O.ke_heat(X, k, 1.5)
    assert st == 0
    assert np.abs(K @ np.ones(8)).max() < 1e-13 * np.abs(K).max()
    g = rng.standard_normal(3)
    u = X @ g
    assert np.isclose(u @ K @ u, V * g @ (k * g), rtol=1e-12)
    assert np.isclose(M.sum(), 1.5 * V, rtol=1e-13)  # consistent mass conserves capacity
    assert np.all(np.linalg.eigvalsh(M) > 0)


def test_patch_test_jittered_mesh():
    """Linear displacement on a jittered 3^3 patch: every interior DOF row of
    K u vanishes (constant stress is reproduced)."""
    prob = meshgen.elastic_cube(3, jitter=0.15)
    prob.mat = np.tile([[10.0, 0.25]], (prob.n_el, 1))
    prob.restrained[:] = 0
    eq = O.assemble_problem(prob)
    A = eq.to_dense()
    G = np.random.default_rng(9).standard_normal((3, 3))
    u = (prob.xyz @ G.T).ravel()
    f = A @ u
    inner = ~meshgen.boundary_mask_hex(3, 3, 3)
    rows = np.concatenate([3 * np.where(inner)[0] + c for c in range(3)])
    assert np.abs(f[rows]).max() < 1e-12 * np.abs(f).max()


# --------------------------------------------------------- pattern + scatter
def _brute_force_global(prob):
    """Dense assembly with the elimination rules of equation.cpp:445-562."""
    n = prob.n_dof
    first, nd = O.dof_numbering(prob.n_nodes, prob.ndof)
    K = np.zeros((n, n))
    F = np.zeros(n)
    rest = prob.restrained.astype(bool)
    for e in range(prob.n_el):
        c = prob.conn[e]
        if prob.dim == 3:
            st, Ke = O.ke_hex8(prob.xyz[c], O.iso_cmat(*prob.mat[e if len(prob.mat) > 1 else 0]))
        else:
            st, Ke = O.ke_quad4(prob.xyz[c], O.plane_stress_cmat(*prob.mat[0]))
        m = np.concatenate([[first[a] + k for k in range(prob.ndof)] for a in c])
        for i, I in enumerate(m):
            for j, J in enumerate(m):
                if rest[J]:
                    if I == J:
                        K[I, I] = 1
                    elif not rest[I]:
                        F[I] -= Ke[i, j] * prob.prescribed[J]
                elif not rest[I]:
                    K[I, J] += Ke[i, j]
    return K, F


@pytest.mark.parametrize("which", ["quad", "hex"])
def test_scatter_equals_brute_force(which):
    if which == "quad":
        prob = meshgen.plane_stress_quad(3)
        prob.prescribed[prob.restrained.astype(bool)] = 0.01 * np.arange(prob.restrained.sum())
    else:
        prob = meshgen.plastic_bar(2)
        prob.mat_kind = capi.MAT_ISO_ELASTIC
        prob.mat = np.array([[5.0, 0.3]])
    eq = O.assemble_problem(prob)
    Kb, Fb = _brute_force_global(prob)
    rp, ci = eq.pattern()
    nzb = np.abs(Kb) > 0
    for i in range(prob.n_dof):
        cols = ci[rp[i]:rp[i + 1]]
        assert list(cols) == sorted(cols) and cols[-1] == i  # sorted, diagonal last
        lower_nz = [j for j in range(i + 1) if nzb[i, j]]
        assert set(lower_nz) <= set(cols)
        if prob.restrained[i]:
            assert list(cols) == [i]
    A = eq.to_dense()
    assert np.allclose(A, Kb, rtol=0, atol=1e-13 * np.abs(Kb).max())
    assert np.allclose(eq.pseudo(), Fb, rtol=0, atol=1e-13 * max(np.abs(Fb).max(), 1))


def test_pattern_closed_form_counts():
    """Config-1 sizes: 64x64 quad4, 130 restrained -> 77,630 lower entries."""
    prob = meshgen.plane_stress_quad(64)
    assert prob.n_dof == 8450 and prob.restrained.sum() == 130
    first, nd = O.dof_numbering(prob.n_nodes, 2)
    eq = O.Equations(prob.n_dof)
    eq.set_restraints(prob.restrained, prob.prescribed)
    eq.specify_mesh(prob.conn, first, nd)
    assert eq.nnz() == 77630


def test_structured_hex_counts_small():
    """n^3 hex8, z=0 clamped: full nnz = sum over free nodes of 3*3*(free
    neighbours incl. self) + restrained DOFs; checked against the oracle."""
    for n in (2, 3, 5):
        prob = meshgen.elastic_cube(n)
        first, nd = O.dof_numbering(prob.n_nodes, 3)
        eq = O.Equations(prob.n_dof)
        eq.set_restraints(prob.restrained, prob.prescribed)
        eq.specify_mesh(prob.conn, first, nd)
        i, j, k = meshgen._node_ijk(n, n, n)
        free = k > 0
        full = 0
        for a in np.where(free)[0]:
            di, dj, dk = np.abs(i - i[a]), np.abs(j - j[a]), np.abs(k - k[a])
            nb = (di <= 1) & (dj <= 1) & (dk <= 1) & free
            full += 9 * nb.sum()
        full += 3 * (~free).sum()
        lower = (full + prob.n_dof) // 2
        assert eq.nnz() == lower


# ------------------------------------------------------------------ solvers
@pytest.mark.parametrize("kind", ["quad", "hex"])
def test_pcg_and_skyline_against_dense(kind):
    prob = meshgen.plane_stress_quad(6) if kind == "quad" else meshgen.elastic_cube(3)
    eq = O.assemble_problem(prob)
    A = eq.to_dense()
    b = eq.rhs()
    x_ref = np.linalg.solve(A, b)
    st, x, it, nrm = eq.pcg(b, eps=1e-12, stop_rule=1)
    assert st == 0 and it > 0
    assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
    st, xd = eq.direct(b)
    assert np.linalg.norm(xd - x_ref) <= 1e-12 * np.linalg.norm(x_ref)
    y = eq.spmv(x_ref)
    assert np.allclose(y, A @ x_ref, rtol=0, atol=1e-13 * np.abs(b).max())


def test_pcg_stop_rules_and_cap():
    prob = meshgen.elastic_cube(3)
    eq = O.assemble_problem(prob)
    b = eq.rhs()
    st, x, it, nrm = eq.pcg(b, eps=1e-6, stop_rule=0)
    assert st == 0
    r = b - eq.to_dense() @ x
    assert nrm < 1e-6
    st, x, it, nrm = eq.pcg(b, eps=1e-14, maxit=3, stop_rule=1)
    assert st == O.NOT_CONVERGED and it == 3
    # x0 equal to the solution: converges at once
    st, x1, it, nrm = eq.pcg(b, eps=1e-10, stop_rule=1)
    st, x2, it2, nrm = eq.pcg(b, x0=x1, eps=1e-8, stop_rule=1)
    assert it2 <= 1


def test_zero_diagonal_reported():
    eq = O.Equations(3)
    eq.specify_mesh(np.array([[0, 1, 2, 0]])[:, :0].reshape(0, 4), [0, 1, 2], [1, 1, 1])
    st, _, _, _ = eq.pcg(np.ones(3), eps=1e-8)
    assert st == O.ZERO_DIAG


# ------------------------------------------------------------- plasticity
def test_j2_elastic_step():
    it, ds, dep, Cep, pl = O.j2_increment(200, 0.3, 0.25, 2.0, np.zeros(6), 0.0,
                                          np.array([1e-4, 0, 0, 0, 0, 0]))
    assert pl == 0 and it == 0 and dep == 0
    assert np.allclose(ds, O.iso_cmat(200, 0.3) @ np.array([1e-4, 0, 0, 0, 0, 0]))


def test_j2_uniaxial_stress_closed_form():
    """Homogeneous uniaxial stress driven in strain increments: each increment
    solves for the lateral strains so that the lateral stresses vanish (a
    Newton loop on the oracle's own tangent).  At eps = 0.5 %, E = 200,
    H = 2, sigma0 = 0.25: sigma = sigma0 + E H/(E+H) (eps - sigma0/E)."""
    E, nu, s0, H = 200.0, 0.3, 0.25, 2.0
    stress = np.zeros(6)
    ep = 0.0
    n_inc = 20
    de_z = 0.005 / n_inc
    lat = 0.0
    for _ in range(n_inc):
        dl = -nu * de_z if lat == 0 else lat
        for _ in range(30):
            de = np.array([dl, dl, de_z, 0, 0, 0])
            it, ds, dep, Cep, pl = O.j2_increment(E, nu, s0, H, stress, ep, de)
            res = (stress + ds)[0]
            if abs(res) < 1e-15:
                break
            dl -= res / (Cep[0, 0] + Cep[0, 1])
        lat = dl
        stress = stress + ds
        ep += dep
    sig = s0 + E * H / (E + H) * (0.005 - s0 / E)
    assert abs(stress[2] - sig) < 1e-9
    assert abs(stress[2] - 0.257425743) < 1e-9


# ------------------------------------------------ plastic and elastic elements
def _jittered_hex(rng, scale=1.0, jit=0.12):
    """A non-affine (warped) hex8 in the reference node order."""
    X = (HEX_NODES + 1) / 2 * scale + jit * scale * rng.uniform(-1, 1, (8, 3))
    return X + rng.standard_normal(3)


def _plastic_call(X, params, ue, isv):
    L = O._plastic_sigs()
    L.orc_plastic_update(O.d(O.f64(X).ravel().copy()), O.d(O.f64(params)), 2, O.d(O.f64(ue).copy()), O.d(isv))
    Ke, Fe = np.zeros(576), np.zeros(24)
    L.orc_plastic_ke_fe(O.d(O.f64(X).ravel().copy()), O.d(O.f64(params)), 2, O.d(isv), O.d(Ke), O.d(Fe))
    K = Ke.reshape(24, 24)
    return np.triu(K) + np.triu(K, 1).T, Fe


def _fe_elastic(X, Cm, ue, order=2):
    L = O.lib()
    L.orc_fe_hex8_elastic.restype = C.c_int
    L.orc_fe_hex8_elastic.argtypes = [O._P_D, O._P_D, C.c_int, O._P_D, O._P_D]
    Fe = np.zeros(24)
    st = L.orc_fe_hex8_elastic(O.d(O.f64(X).ravel().copy()), O.d(O.f64(Cm).ravel().copy()), order,
                               O.d(O.f64(ue).copy()), O.d(Fe))
    return st, Fe


@pytest.mark.parametrize("order", [2, 3])
def test_fe_hex8_elastic_is_ke_times_ue(order):
    """orc_fe_hex8_elastic (addToFe with stress = C strain) on warped bricks
    equals K_e u_e with the pinned K_e (rigid modes, uniform-strain energy,
    patch test) for random u_e; a wrong B row, shear factor or stress
    component in the force path fails it."""
    rng = np.random.default_rng(31)
    for _ in range(4):
        X = _jittered_hex(rng, scale=rng.uniform(0.2, 2.0))
        Cm = O.iso_cmat(rng.uniform(1, 100), rng.uniform(0.15, 0.4))
        st, K = O.ke_hex8(X, Cm, order)
        assert st == 0
        ue = rng.standard_normal(24)
        st, Fe = _fe_elastic(X, Cm, ue, order)
        assert st == 0
        assert np.abs(Fe - K @ ue).max() <= 1e-13 * np.abs(K).max() * np.abs(ue).max()


def test_plastic_element_elastic_increment_warped():
    """A purely elastic increment with shear on a warped hex8, from a nonzero
    stress below yield: the plastic element's "KI" tangent equals the pinned
    elastic K_e and its "F" force equals K_e u_e (the force of the stress
    increment), at every Gauss point the return map stays elastic."""
    rng = np.random.default_rng(32)
    E, nu, s0, H = 200.0, 0.3, 0.25, 2.0
    params = np.array([E, nu, s0, H])
    for _ in range(4):
        X = _jittered_hex(rng)
        ue = 1e-5 * rng.standard_normal(24)
        isv = np.zeros((8, O.PISV))
        isv[:, :6] = 0.02 * rng.standard_normal(6)  # prior stress, well inside the yield surface
        isv[:, 6] = 0.001
        Kp, Fp = _plastic_call(X, params, ue, isv)
        _, K = O.ke_hex8(X, O.iso_cmat(E, nu))
        assert np.all(isv[:, 14] == 0)
        assert np.abs(Kp - K).max() <= 1e-13 * np.abs(K).max()
        assert np.abs(Fp - K @ ue).max() <= 1e-13 * np.abs(K).max() * np.abs(ue).max()


def test_plastic_pure_shear_path_closed_form():
    """Homogeneous simple shear u = gamma*y on an affine (sheared, scaled)
    hex8, driven in increments past yield, "V" + "KI"/"F" + "U" per increment.

    Closed form of the reference's return map (getStressIncrement version
    2.0, plastic.cpp:368-465) in pure shear tau = sigma_xy, J2 with linear
    hardening: the Hill function is tau^2, the yield function
    3 tau^2 / sigma_y^2 - 1 with sigma_y = sigma0 + H epbar, the flow vector
    (0,0,0, 2 tau, 0,0).  An increment dgamma is elastic (tau += G dgamma)
    when the trial 3 (tau + G dgamma)^2 <= sigma_y^2; otherwise the whole
    increment is plastic and one return-map pass gives
      tau += G dgamma H / (3G + H)            (the J2 elastoplastic shear modulus)
      epbar += sqrt(3) tau_new G dgamma / ((H + 3G) tau_trial)
    (stop test met after that pass), and the tangent's shear entry is
    C_ep[3][3] = G H / (3G + H) with every other entry elastic.  The element
    force of the increment obeys the virial identity sum_a x_a (x) F_a =
    V dsigma (exact for an affine brick).  A dropped shear factor in the Hill
    function or the flow vector, tensor instead of engineering shear, or a
    wrong hardening modulus changes these numbers."""
    E, nu, s0, H = 200.0, 0.3, 0.25, 2.0
    G = E / (2 * (1 + nu))
    params = np.array([E, nu, s0, H])
    rng = np.random.default_rng(33)
    X, V = _affine_hex(rng, scale=0.7)
    isv = np.zeros((8, O.PISV))
    tau, ep = 0.0, 0.0
    g_y = s0 / math.sqrt(3) / G
    steps = [0.6 * g_y, 0.3 * g_y, 0.5 * g_y, 0.5 * g_y, 1.0 * g_y, 2.0 * g_y, 0.02 * g_y]
    flags = []
    for dg in steps:
        ue = np.zeros((8, 3))
        ue[:, 0] = dg * X[:, 1]
        Kp, Fp = _plastic_call(X, params, ue.ravel(), isv)
        tr = tau + G * dg
        plastic = 3 * tr ** 2 > (s0 + H * ep) ** 2
        flags.append(bool(plastic))
        if plastic:
            tau_new = tau + G * dg * H / (3 * G + H)
            dep = math.sqrt(3) * tau_new * G * dg / ((H + 3 * G) * tr)
        else:
            tau_new, dep = tr, 0.0
        dtau = tau_new - tau
        for ip in range(8):
            r = isv[ip]
            assert r[14] == (1 if plastic else 0)
            assert abs(r[7 + 3] - dtau) <= 1e-12 * abs(dtau)
            assert np.abs(r[[7, 8, 9, 11, 12]]).max() <= 1e-12 * abs(dtau)
            assert abs(r[13] - dep) <= 1e-12 * max(dep, 1e-30)
            Cep = r[16:52].reshape(6, 6)
            Ce = O.iso_cmat(E, nu)
            want = Ce.copy()
            if plastic:
                want[3, 3] = G * H / (3 * G + H)
            assert np.abs(Cep - want).max() <= 1e-12 * E
        # virial identity on the force of the stress increment
        Fa = Fp.reshape(8, 3)
        M = Fa.T @ X  # M[i, j] = sum_a F_a(i) x_a(j)
        dsig = np.zeros((3, 3))
        dsig[0, 1] = dsig[1, 0] = dtau
        assert np.abs(M - V * dsig).max() <= 1e-12 * V * abs(dtau)
        L = O._plastic_sigs()
        L.orc_plastic_commit(O.d(isv), 8)
        tau, ep = tau_new, ep + dep
        assert np.allclose(isv[:, 3], tau, rtol=1e-12, atol=0)
    # the first crossing makes its whole increment plastic and leaves the
    # state inside the surface (tau ~ 0.9 tau_y); later large steps cross
    # again, the last small one stays elastic
    assert flags == [False, False, True, True, True, True, False]


def test_heat_consistent_capacity_closed_form():
    """Consistent capacity matrix of an affine brick (parallelepiped):
    M_ab = c V / 216 * {8, 4, 2, 1} for nodes a, b that differ in 0, 1, 2, 3
    reference coordinates (the tensor product of the 1D h/6 [2 1; 1 2]).  A
    lumped or row-summed M, a missing detJ or a wrong shape function fails it."""
    rng = np.random.default_rng(34)
    for _ in range(3):
        X, V = _affine_hex(rng, scale=rng.uniform(0.3, 2.0))
        cap = rng.uniform(0.5, 3.0)
        st, K, M = O.ke_heat(X, np.array([1.0, 1.0, 1.0]), cap)
        assert st == 0
        ndiff = (HEX_NODES[:, None, :] != HEX_NODES[None, :, :]).sum(axis=2)
        want = cap * V / 216.0 * np.array([8.0, 4.0, 2.0, 1.0])[ndiff]
        assert np.abs(M - want).max() <= 1e-14 * cap * V


def test_j2_algorithmic_derivative_matches_finite_differences():
    """orc_j2_stress_increment_alg's Dalg is the derivative of the return map
    it executes: central differences of dstress over dstrain (same branch and
    iteration count) agree to 1e-7 of |D|.  In a proportional (uniaxial)
    state it is symmetric and its symmetric part is what the assembly uses."""
    E, nu, s0, H = 200.0, 0.3, 0.25, 2.0
    rng = np.random.default_rng(35)
    h = 1e-8
    for _ in range(6):
        stress = rng.standard_normal(6) * 0.05
        de = rng.standard_normal(6) * 2e-3
        it, ds, dep, Cep, D, pl = O.j2_increment_alg(E, nu, s0, H, stress, 0.001, de)
        assert pl == 1
        it2, ds2, dep2, Cep2, pl2 = O.j2_increment(E, nu, s0, H, stress, 0.001, de)
        assert np.array_equal(ds, ds2) and np.array_equal(Cep, Cep2)  # same update, same reference tangent
        Dfd = np.zeros((6, 6))
        for j in range(6):
            e1, e2 = de.copy(), de.copy()
            e1[j] += h
            e2[j] -= h
            r1 = O.j2_increment_alg(E, nu, s0, H, stress, 0.001, e1)
            r2 = O.j2_increment_alg(E, nu, s0, H, stress, 0.001, e2)
            assert r1[0] == r2[0] == it
            Dfd[:, j] = (r1[1] - r2[1]) / (2 * h)
        assert np.abs(D - Dfd).max() <= 1e-7 * np.abs(D).max()
    # elastic step: D = C
    it, ds, dep, Cep, D, pl = O.j2_increment_alg(E, nu, s0, H, np.zeros(6), 0.0, np.array([1e-4, 0, 0, 0, 0, 0]))
    assert pl == 0 and np.array_equal(D, O.iso_cmat(E, nu))
    # uniaxial trial state: symmetric derivative
    sig = np.array([0, 0, 0.2, 0, 0, 0.0])
    it, ds, dep, Cep, D, pl = O.j2_increment_alg(E, nu, s0, H, sig, 0.0, np.array([-3e-4, -3e-4, 1e-3, 0, 0, 0]))
    assert pl == 1 and np.abs(D - D.T).max() <= 1e-12 * np.abs(D).max()


def test_plastic_driver_all_plastic_path_closed_form():
    """PlasticityModel driver (oracle plastic_analysis) on the rollers bar,
    2^3 elements, in four increments of 0.3% axial strain (each 2.4x the
    yield strain), so every Gauss point takes the plastic branch in every
    increment.  For a uniaxial trial state the flow vector's direction does
    not depend on its size and the reference's one-pass update is linear in
    the strain increment, with the J2 continuum modulus: each increment adds
    E H / (E + H) times its axial strain to sigma_zz and nothing laterally.
    So sigma_zz = E H/(E+H) * 1.2% at every point (the reference's update
    does not carry sigma0 into a first plastic increment from zero stress).
    With the algorithmic tangent Newton reaches it in a few iterations; the
    reference tangent needs hundreds (it converges linearly)."""
    prob = meshgen.plastic_bar(2, strain=0.012, increments=4)
    u, isv, hist = O.plastic_analysis(prob, 4, 1e-9, 300, tangent=O.TANGENT_ALGORITHMIC)
    Et = 200.0 * 2.0 / 202.0
    assert all(h[0] <= 8 and h[1] < 2e-9 for h in hist), hist
    assert np.abs(isv[:, :, 2] - Et * 0.012).max() <= 1e-8
    assert np.abs(isv[:, :, [0, 1, 3, 4, 5]]).max() <= 1e-7
    assert np.all(isv[:, :, 14] == 1)
    # the top face moved by the prescribed 1.2 %
    top = np.isclose(prob.xyz[:, 2], 1.0)
    assert np.allclose(u[3 * np.where(top)[0] + 2], 0.012, rtol=0, atol=1e-15)


@pytest.mark.parametrize("n", [3, 4])
def test_parcg_two_threads_matches_serial(n):
    """ParCGMatrix::solve restated with two OpenMP threads: the same solution
    as the serial sparse.cpp PCG and a dense solve, iterations within one
    (only the grouping of partial sums differs); at n = 4 (375 DOF) the
    L/2 split cuts a node."""
    prob = meshgen.elastic_cube(n)
    eq = O.assemble_problem(prob)
    b = eq.rhs()
    st1, x1, it1, _ = eq.pcg(b, eps=1e-12, stop_rule=0, maxit=10000)
    st2, x2, it2, _ = eq.pcg_parcg(b, eps=1e-12, maxit=10000)
    assert st1 == 0 and st2 == 0 and abs(it1 - it2) <= 1
    A = eq.to_dense()
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(x2 - xd) <= 1e-9 * np.linalg.norm(xd)
    assert np.linalg.norm(x2 - x1) <= 1e-9 * np.linalg.norm(x1)
    # the two-thread product equals the serial one
    y = eq.spmv(b)
    assert np.allclose(A @ b, y, rtol=0, atol=1e-12 * np.abs(y).max())
