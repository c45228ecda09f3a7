"""ctypes binding of the oracle (oracle/beta_oracle.c).

TEST INFRASTRUCTURE ONLY: the parity checker.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module; the product path (betanoox_amd/) never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libbeta_oracle.so")
# the CPU baseline's -O3 -march=native build (make -C oracle native), used
# when ORACLE_NATIVE=1 is set before the first call
NATIVE_PATH = os.path.join(HERE, "_build", "native", "libbeta_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbeta_ref.so")

_P_D = C.POINTER(C.c_double)
_P_I = C.POINTER(C.c_int)
_P_U8 = C.POINTER(C.c_ubyte)

OK, ERR_ARG, NOT_CONVERGED, ZERO_DIAG, NEG_DIAG, NEG_DETJ = 0, -1, -4, -5, -6, -7

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    path = NATIVE_PATH if os.environ.get("ORACLE_NATIVE") == "1" else LIB_PATH
    if not os.path.exists(path):
        raise OSError("oracle not built: make -C oracle (native: make -C oracle native)")
    L = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "orc_gauss": (None, [C.c_int, _P_D, _P_D]),
        "orc_hex_quad_points": (C.c_int, [C.c_int, _P_D, _P_D, _P_D, _P_D]),
        "orc_quad_quad_points": (C.c_int, [C.c_int, _P_D, _P_D, _P_D]),
        "orc_shape_hex8": (None, [_P_D, C.c_double, C.c_double, C.c_double]),
        "orc_deriv_hex8": (None, [_P_D, _P_D, _P_D, C.c_double, C.c_double, C.c_double]),
        "orc_shape_quad4": (None, [_P_D, C.c_double, C.c_double]),
        "orc_deriv_quad4": (None, [_P_D, _P_D, C.c_double, C.c_double]),
        "orc_jacobian3d": (C.c_double, [C.c_int, _P_D, _P_D, _P_D, _P_D, _P_D, _P_D, _P_D, _P_D, _P_D]),
        "orc_ludcmp": (C.c_int, [C.c_int, _P_D, _P_I, _P_D]),
        "orc_lubksb": (None, [C.c_int, _P_D, _P_I, _P_D]),
        "orc_invert": (C.c_int, [C.c_int, _P_D]),
        "orc_elastic_cmat": (C.c_int, [_P_D, _P_D]),
        "orc_elastic_iso_cmat": (C.c_int, [C.c_double, C.c_double, _P_D]),
        "orc_plane_stress_cmat": (None, [C.c_double, C.c_double, C.c_double, C.c_double, _P_D]),
        "orc_ke_hex8_elastic": (C.c_int, [_P_D, _P_D, C.c_int, _P_D]),
        "orc_ke_quad4_elastic": (C.c_int, [_P_D, _P_D, C.c_int, C.c_double, _P_D]),
        "orc_ke_hex8_heat": (C.c_int, [_P_D, _P_D, C.c_double, C.c_int, _P_D, _P_D]),
        "orc_fe_hex8_elastic": (C.c_int, [_P_D, _P_D, C.c_int, _P_D, _P_D]),
        "orc_eq_create": (vp, [C.c_int]),
        "orc_eq_destroy": (None, [vp]),
        "orc_eq_set_restraint": (None, [vp, C.c_int, C.c_double]),
        "orc_eq_specify": (None, [vp, _P_I, C.c_int]),
        "orc_eq_allocate": (None, [vp]),
        "orc_eq_nnz": (C.c_longlong, [vp]),
        "orc_eq_get_pattern": (None, [vp, _P_I, _P_I]),
        "orc_eq_get_values": (None, [vp, _P_D]),
        "orc_eq_set_values": (None, [vp, _P_D]),
        "orc_eq_zero": (None, [vp]),
        "orc_eq_add_matrix": (C.c_int, [vp, _P_D, _P_I, C.c_int, _P_D]),
        "orc_eq_add_load": (C.c_int, [vp, C.c_int, C.c_double]),
        "orc_eq_build_rhs": (None, [vp, _P_D]),
        "orc_eq_get_pseudo": (None, [vp, _P_D]),
        "orc_eq_set_replace_zero_diag": (None, [vp, C.c_double]),
        "orc_eq_spmv": (None, [vp, _P_D, _P_D]),
        "orc_eq_check_fix_diag": (C.c_int, [vp]),
        "orc_eq_pcg": (C.c_int, [vp, _P_D, _P_D, C.c_double, C.c_int, C.c_int, _P_I, _P_D]),
        "orc_eq_skyline_solve": (C.c_int, [vp, _P_D]),
        "orc_dof_list": (C.c_int, [C.c_int, _P_I, _P_I, _P_I, _P_I]),
        "orc_specify_mesh": (None, [vp, C.c_int, C.c_int, _P_I, _P_I, _P_I]),
        "orc_assemble_elastic3d": (C.c_int, [vp, C.c_int, _P_I, _P_D, _P_I, _P_I, _P_D, _P_I, C.c_int]),
        "orc_assemble_elastic2d": (C.c_int, [vp, C.c_int, _P_I, _P_D, _P_I, _P_I, _P_D, _P_I, C.c_int, C.c_double]),
        "orc_assemble_heat_transient": (C.c_int, [vp, C.c_int, _P_I, _P_D, _P_I, _P_I, _P_D, _P_I, C.c_int,
                                                  C.c_double, C.c_double, _P_D]),
        "orc_j2_stress_increment": (C.c_int, [_P_D, _P_D, C.c_double, _P_D, _P_D, _P_D, _P_D, _P_I]),
        "orc_eq_pcg_parcg": (C.c_int, [vp, _P_D, _P_D, C.c_double, C.c_int, _P_I, _P_D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def d(a):
    return None if a is None else a.ctypes.data_as(_P_D)


def i32(a):
    return None if a is None else a.ctypes.data_as(_P_I)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def ai32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ------------------------------------------------------------ small helpers
def gauss(k):
    c, w = np.zeros(k), np.zeros(k)
    lib().orc_gauss(k, d(c), d(w))
    return c, w


def hex_points(order):
    n = order ** 3
    a = [np.zeros(n) for _ in range(4)]
    lib().orc_hex_quad_points(order, *[d(x) for x in a])
    return a


def iso_cmat(E, nu):
    C6 = np.zeros(36)
    lib().orc_elastic_iso_cmat(E, nu, d(C6))
    return C6.reshape(6, 6)


def plane_stress_cmat(e11, e22, nu12, G12):
    C3 = np.zeros(9)
    lib().orc_plane_stress_cmat(e11, e22, nu12, G12, d(C3))
    return C3.reshape(3, 3)


def invert(a):
    a = f64(a).copy()
    st = lib().orc_invert(a.shape[0], d(a))
    return st, a


def ke_hex8(xyz8, Cm, order=2):
    Ke = np.zeros(576)
    st = lib().orc_ke_hex8_elastic(d(f64(xyz8).ravel().copy()), d(f64(Cm).ravel().copy()), order, d(Ke))
    K = Ke.reshape(24, 24)
    return st, np.triu(K) + np.triu(K, 1).T


def ke_quad4(xy4, Cm, order=2, factor=1.0):
    Ke = np.zeros(64)
    st = lib().orc_ke_quad4_elastic(d(f64(xy4).ravel().copy()), d(f64(Cm).ravel().copy()), order, factor, d(Ke))
    K = Ke.reshape(8, 8)
    return st, np.triu(K) + np.triu(K, 1).T


def ke_heat(xyz8, k3, cap, order=2):
    Ke, Me = np.zeros(64), np.zeros(64)
    k9 = np.diag(k3).ravel().astype(np.float64)
    st = lib().orc_ke_hex8_heat(d(f64(xyz8).ravel().copy()), d(k9), cap, order, d(Ke), d(Me))
    K, M = Ke.reshape(8, 8), Me.reshape(8, 8)
    return st, np.triu(K) + np.triu(K, 1).T, np.triu(M) + np.triu(M, 1).T


def j2_increment_alg(E, nu, sigma0, H, stress, eps_bar, dstrain):
    """As j2_increment, plus the exact derivative d(dstress)/d(dstrain) of
    the return map as executed (orc_j2_stress_increment_alg)."""
    L = lib()
    L.orc_j2_stress_increment_alg.restype = C.c_int
    L.orc_j2_stress_increment_alg.argtypes = [_P_D, _P_D, C.c_double, _P_D, _P_D, _P_D, _P_D, _P_D,
                                              C.POINTER(C.c_int)]
    p = np.array([E, nu, sigma0, H], dtype=np.float64)
    ds, dep, Cep, D, pl = np.zeros(6), np.zeros(1), np.zeros(36), np.zeros(36), C.c_int()
    it = L.orc_j2_stress_increment_alg(d(p), d(f64(stress).copy()), eps_bar, d(f64(dstrain).copy()), d(ds),
                                       d(dep), d(Cep), d(D), C.byref(pl))
    return it, ds, dep[0], Cep.reshape(6, 6), D.reshape(6, 6), pl.value


def j2_increment(E, nu, sigma0, H, stress, eps_bar, dstrain):
    p = np.array([E, nu, sigma0, H], dtype=np.float64)
    ds, dep, Cep, pl = np.zeros(6), np.zeros(1), np.zeros(36), C.c_int()
    it = lib().orc_j2_stress_increment(d(p), d(f64(stress).copy()), eps_bar, d(f64(dstrain).copy()), d(ds),
                                       d(dep), d(Cep), C.byref(pl))
    return it, ds, dep[0], Cep.reshape(6, 6), pl.value


# ------------------------------------------------------------ Equations
class Equations:
    """The reference's Equations + SymmetricSparseMatrix (lower storage)."""

    def __init__(self, n_dof):
        self.L = lib()
        self.h = self.L.orc_eq_create(n_dof)
        self.n = n_dof

    def __del__(self):
        try:
            if self.h:
                self.L.orc_eq_destroy(self.h)
        except Exception:
            pass

    def set_restraints(self, restrained, values):
        for i in np.where(np.asarray(restrained) != 0)[0]:
            self.L.orc_eq_set_restraint(self.h, int(i), float(values[i]))

    def specify_mesh(self, conn, dof_first, ndof):
        conn = ai32(conn)
        self.L.orc_specify_mesh(self.h, conn.shape[0], conn.shape[1], i32(conn), i32(ai32(dof_first)),
                                i32(ai32(ndof)))

    def nnz(self):
        return int(self.L.orc_eq_nnz(self.h))

    def pattern(self):
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(max(self.nnz(), 1), np.int32)
        self.L.orc_eq_get_pattern(self.h, i32(rp), i32(ci))
        return rp, ci[: self.nnz()]

    def values(self):
        v = np.zeros(max(self.nnz(), 1))
        self.L.orc_eq_get_values(self.h, d(v))
        return v[: self.nnz()]

    def set_values(self, v):
        self.L.orc_eq_set_values(self.h, d(f64(v)))

    def add_load(self, load):
        for i in np.nonzero(load)[0]:
            st = self.L.orc_eq_add_load(self.h, int(i), float(load[i]))
            if st != OK:
                raise ValueError("load on restrained DOF %d" % i)

    def rhs(self):
        b = np.zeros(self.n)
        self.L.orc_eq_build_rhs(self.h, d(b))
        return b

    def pseudo(self):
        f = np.zeros(self.n)
        self.L.orc_eq_get_pseudo(self.h, d(f))
        return f

    def spmv(self, x):
        y = np.zeros(self.n)
        self.L.orc_eq_spmv(self.h, d(f64(x)), d(y))
        return y

    def pcg(self, b, x0=None, eps=1e-8, maxit=10000, stop_rule=1):
        bb = f64(b).copy()
        it, nrm = C.c_int(), C.c_double()
        st = self.L.orc_eq_pcg(self.h, d(bb), d(None if x0 is None else f64(x0)), eps, maxit, stop_rule,
                               C.byref(it), C.byref(nrm))
        return st, bb, it.value, nrm.value

    def pcg_parcg(self, b, x0=None, eps=1e-8, maxit=10000):
        """ParCGMatrix::solve: two threads, ||r||/||x|| stop rule."""
        bb = f64(b).copy()
        it, nrm = C.c_int(), C.c_double()
        st = self.L.orc_eq_pcg_parcg(self.h, d(bb), d(None if x0 is None else f64(x0)), eps, maxit,
                                     C.byref(it), C.byref(nrm))
        return st, bb, it.value, nrm.value

    def direct(self, b):
        bb = f64(b).copy()
        st = self.L.orc_eq_skyline_solve(self.h, d(bb))
        return st, bb

    def to_dense(self):
        rp, ci = self.pattern()
        v = self.values()
        A = np.zeros((self.n, self.n))
        for i in range(self.n):
            for k in range(rp[i], rp[i + 1]):
                A[i, ci[k]] = v[k]
                A[ci[k], i] = v[k]
        return A


def dof_numbering(n_nodes, ndof):
    nd = np.full(n_nodes, ndof, np.int32)
    first = np.concatenate([[0], np.cumsum(nd)[:-1]]).astype(np.int32)
    return first, nd


def assemble_problem(prob, elem_group=None, cmats=None):
    """Pattern + assembly + loads exactly as the reference's serial path.
    Returns the Equations object (K, pseudo-load, load set)."""
    from betanoox_amd import capi  # constants only
    first, nd = dof_numbering(prob.n_nodes, prob.ndof)
    eq = Equations(prob.n_dof)
    eq.set_restraints(prob.restrained, prob.prescribed)
    eq.specify_mesh(prob.conn, first, nd)
    L = lib()
    conn = ai32(prob.conn)
    xyz = f64(prob.xyz)
    if elem_group is None and len(prob.mat) == 1:
        elem_group = np.zeros(prob.n_el, np.int32)
    g = None if elem_group is None else ai32(elem_group)
    if prob.mat_kind in (capi.MAT_ISO_ELASTIC, capi.MAT_J2):
        if cmats is None:
            cmats = np.stack([iso_cmat(m[0], m[1]).ravel() for m in prob.mat])
        st = L.orc_assemble_elastic3d(eq.h, conn.shape[0], i32(conn), d(xyz), i32(first), i32(nd),
                                      d(f64(cmats)), i32(g), prob.order)
    elif prob.mat_kind == capi.MAT_CMAT3D:
        st = L.orc_assemble_elastic3d(eq.h, conn.shape[0], i32(conn), d(xyz), i32(first), i32(nd),
                                      d(f64(prob.mat)), i32(g), prob.order)
    elif prob.mat_kind == capi.MAT_PLANE_STRESS:
        cm = np.stack([plane_stress_cmat(*m).ravel() for m in prob.mat])
        st = L.orc_assemble_elastic2d(eq.h, conn.shape[0], i32(conn), d(xyz), i32(first), i32(nd),
                                      d(cm), i32(g), prob.order, prob.factor)
    else:
        raise ValueError("use assemble_heat for heat problems")
    if st != OK:
        raise RuntimeError("oracle assembly failed: %d" % st)
    eq.add_load(prob.load)
    return eq


def assemble_heat(prob, alpha, dt, u_s):
    first, nd = dof_numbering(prob.n_nodes, 1)
    eq = Equations(prob.n_dof)
    eq.set_restraints(prob.restrained, prob.prescribed)
    eq.specify_mesh(prob.conn, first, nd)
    conn = ai32(prob.conn)
    g = np.zeros(prob.n_el, np.int32) if len(prob.mat) == 1 else None
    st = lib().orc_assemble_heat_transient(eq.h, conn.shape[0], i32(conn), d(f64(prob.xyz)), i32(first),
                                           i32(nd), d(f64(prob.mat)), i32(g), prob.order, alpha, dt,
                                           d(f64(u_s)))
    if st != OK:
        raise RuntimeError("oracle heat assembly failed: %d" % st)
    return eq


# ------------------------------------------------------------ oracle/_ref
_ref = None


def ref_lib():
    """The reference's own shape.cpp + SparseMatrixUtility.cpp, built by
    `make -C oracle ref` (C++ symbols are called by their mangled names)."""
    global _ref
    if _ref is not None:
        return _ref
    if not os.path.exists(REF_PATH):
        return None
    R = C.CDLL(REF_PATH)
    R._Z9shapeHex8Pdddd.argtypes = [_P_D, C.c_double, C.c_double, C.c_double]
    R._Z9derivHex8PdS_S_ddd.argtypes = [_P_D, _P_D, _P_D, C.c_double, C.c_double, C.c_double]
    R._Z10shapeQuad4Pddd.argtypes = [_P_D, C.c_double, C.c_double]
    R._Z10derivQuad4PdS_dd.argtypes = [_P_D, _P_D, C.c_double, C.c_double]
    ppi, ppd = C.POINTER(_P_I), C.POINTER(_P_D)
    # SparseMatrixUtility.cpp:11-85 (int*& / double*& arguments are pointers to pointers)
    R._Z21writeCSRMatrixInASCIIPcRPiS1_RPdii.argtypes = [C.c_char_p, ppi, ppi, ppd, C.c_int, C.c_int]
    R._Z22writeCSRMatrixInBinaryPcRPiS1_RPdii.argtypes = [C.c_char_p, ppi, ppi, ppd, C.c_int, C.c_int]
    R._Z20readCSRMatrixInASCIIPcRPiS1_RPdRi.argtypes = [C.c_char_p, ppi, ppi, ppd, _P_I]
    R._Z21readCSRMatrixInBinaryPcRPiS1_RPdRiS4_.argtypes = [C.c_char_p, ppi, ppi, ppd, _P_I, _P_I]
    R._Z14readLoadVectorPcPdi.argtypes = [C.c_char_p, _P_D, C.c_int]
    for f in (R._Z21writeCSRMatrixInASCIIPcRPiS1_RPdii, R._Z22writeCSRMatrixInBinaryPcRPiS1_RPdii,
              R._Z20readCSRMatrixInASCIIPcRPiS1_RPdRi, R._Z21readCSRMatrixInBinaryPcRPiS1_RPdRiS4_,
              R._Z14readLoadVectorPcPdi):
        f.restype = None
    _ref = R
    return R


def ref_write_csr(path, binary, ia, ja, a):
    """The reference's writeCSRMatrixInASCII / writeCSRMatrixInBinary
    (SparseMatrixUtility.cpp:11-38) on the given arrays."""
    R = ref_lib()
    ia, ja, a = ai32(ia), ai32(ja), f64(a)
    pia, pja, pa = i32(ia), i32(ja), d(a)
    fn = R._Z22writeCSRMatrixInBinaryPcRPiS1_RPdii if binary else R._Z21writeCSRMatrixInASCIIPcRPiS1_RPdii
    fn(str(path).encode(), C.byref(pia), C.byref(pja), C.byref(pa), ia.shape[0] - 1, ja.shape[0])


def ref_read_csr(path, binary):
    """The reference's readCSRMatrixInASCII / readCSRMatrixInBinary
    (SparseMatrixUtility.cpp:40-78); returns (ia, ja, a) as numpy copies."""
    R = ref_lib()
    pia, pja, pa = _P_I(), _P_I(), _P_D()
    n = C.c_int()
    if binary:
        nc = C.c_int()
        R._Z21readCSRMatrixInBinaryPcRPiS1_RPdRiS4_(str(path).encode(), C.byref(pia), C.byref(pja), C.byref(pa),
                                                     C.byref(n), C.byref(nc))
        m = nc.value
    else:
        R._Z20readCSRMatrixInASCIIPcRPiS1_RPdRi(str(path).encode(), C.byref(pia), C.byref(pja), C.byref(pa),
                                                C.byref(n))
        m = None
    ia = np.ctypeslib.as_array(pia, shape=(n.value + 1,)).copy()
    if m is None:
        m = int(ia[-1] - ia[0])
    ja = np.ctypeslib.as_array(pja, shape=(m,)).copy()
    a = np.ctypeslib.as_array(pa, shape=(m,)).copy()
    return ia, ja, a


def ref_read_load_vector(path, n):
    """The reference's readLoadVector (SparseMatrixUtility.cpp:79-85)."""
    v = np.zeros(n)
    ref_lib()._Z14readLoadVectorPcPdi(str(path).encode(), d(v), n)
    return v


# ------------------------------------------------------------ plasticity
PISV = 52


def _plastic_sigs():
    L = lib()
    if getattr(L, "_plastic_sigs", False):
        return L
    L.orc_plastic_update.restype = C.c_int
    L.orc_plastic_update.argtypes = [_P_D, _P_D, C.c_int, _P_D, _P_D]
    L.orc_plastic_update_tangent.restype = C.c_int
    L.orc_plastic_update_tangent.argtypes = [_P_D, _P_D, C.c_int, _P_D, _P_D, C.c_int]
    L.orc_plastic_ke_fe.restype = C.c_int
    L.orc_plastic_ke_fe.argtypes = [_P_D, _P_D, C.c_int, _P_D, _P_D, _P_D]
    L.orc_plastic_commit.restype = None
    L.orc_plastic_commit.argtypes = [_P_D, C.c_int]
    L._plastic_sigs = True
    return L


TANGENT_CONTINUUM = 0
TANGENT_ALGORITHMIC = 1


def plastic_analysis(prob, increments, max_allowable_residual, max_iterations, solve_eps=1e-12,
                     presc_list=None, tangent=TANGENT_CONTINUUM):
    """PlasticityModel::DoPlasticAnalysis (PlasticityModel.cpp:142-398) over
    the oracle's sparse system: per increment, prescribed increment values on
    restrained DOFs, then "V", assembleKandF ("KI"), getGlobalForces ("F"),
    residual = load - resultant on free DOFs, solve, add on free DOFs; stop
    when max|residual| < 2*max_allowable_residual (checked from the second
    pass on) or after max_iterations extra solves; then "U".  Returns the
    accumulated displacement, the Gauss-point records and per-increment
    (newton solves, max residual, plastic points, largest return-map
    iteration count of the committed records).  presc_list optionally gives the
    prescribed increment vector of every increment (default: prob.prescribed).
    tangent selects the reference's continuum C_ep (default) or the symmetric
    part of the return map's derivative on every pass after the first of an
    increment (the first pass, with only the prescribed increment applied,
    takes the continuum tangent)."""
    L = _plastic_sigs()
    n = prob.n_dof
    order = prob.order
    nip = order ** 3
    p = f64(prob.mat[0][:4])
    first, nd = dof_numbering(prob.n_nodes, 3)
    eq = Equations(n)
    eq.set_restraints(prob.restrained, prob.prescribed)
    eq.specify_mesh(prob.conn, first, nd)
    rest = prob.restrained.astype(bool)
    conn = ai32(prob.conn)
    maps = (3 * conn[:, :, None] + np.arange(3)[None, None, :]).reshape(conn.shape[0], 24).astype(np.int32)
    xyz_e = [f64(prob.xyz[c]).ravel().copy() for c in conn]
    isv = np.zeros((prob.n_el, nip, PISV))
    utot = np.zeros(n)
    hist = []
    load = prob.load
    Ke = np.zeros(576)
    Fe = np.zeros(24)
    for inc in range(increments):
        presc = prob.prescribed if presc_list is None else presc_list[inc]
        du = np.zeros(n)
        du[rest] = presc[rest]
        it = 0
        while True:
            for e in range(prob.n_el):
                rec = isv[e]
                L.orc_plastic_update_tangent(d(xyz_e[e]), d(p), order, d(f64(du[maps[e]])), d(rec),
                                             tangent if it > 0 else TANGENT_CONTINUUM)
            L.orc_eq_zero(eq.h)
            res = np.zeros(n)
            for e in range(prob.n_el):
                L.orc_plastic_ke_fe(d(xyz_e[e]), d(p), order, d(isv[e]), d(Ke), d(Fe))
                st = L.orc_eq_add_matrix(eq.h, d(Ke), i32(maps[e]), 24, None)
                if st != OK:
                    raise RuntimeError("plastic scatter failed")
                np.add.at(res, maps[e], Fe)
            resid = np.where(rest, 0.0, load - res)
            maxres = np.abs(resid).max()
            if it > 0 and maxres < 2 * max_allowable_residual:
                break
            if it > max_iterations:
                break
            st, dsol, _, _ = eq.pcg(resid, eps=solve_eps, stop_rule=1, maxit=100000)
            if st != OK:
                raise RuntimeError("oracle PCG failed in Newton: %d" % st)
            du[~rest] += dsol[~rest]
            it += 1
        for e in range(prob.n_el):
            L.orc_plastic_commit(d(isv[e]), nip)
        utot += du
        hist.append((it, maxres, int((isv[:, :, 14] != 0).sum()), int(isv[:, :, 15].max())))
    return utot, isv, hist


def plastic_tangent_system(prob, du, tangent=TANGENT_CONTINUUM, isv=None):
    """One "V" (return map of the increment displacement du from the records
    isv, zero when None) and the "KI" assembly of PlasticElement::update
    (plasticElement.cpp:39-250) over the oracle's sparse system; returns the
    Equations object and the updated records."""
    L = _plastic_sigs()
    order = prob.order
    nip = order ** 3
    p = f64(prob.mat[0][:4])
    first, nd = dof_numbering(prob.n_nodes, 3)
    eq = Equations(prob.n_dof)
    eq.set_restraints(prob.restrained, prob.prescribed)
    eq.specify_mesh(prob.conn, first, nd)
    conn = ai32(prob.conn)
    maps = (3 * conn[:, :, None] + np.arange(3)[None, None, :]).reshape(conn.shape[0], 24).astype(np.int32)
    if isv is None:
        isv = np.zeros((prob.n_el, nip, PISV))
    Ke, Fe = np.zeros(576), np.zeros(24)
    for e in range(prob.n_el):
        xe = f64(prob.xyz[conn[e]]).ravel().copy()
        L.orc_plastic_update_tangent(d(xe), d(p), order, d(f64(du[maps[e]])), d(isv[e]), tangent)
        L.orc_plastic_ke_fe(d(xe), d(p), order, d(isv[e]), d(Ke), d(Fe))
        if L.orc_eq_add_matrix(eq.h, d(Ke), i32(maps[e]), 24, None) != OK:
            raise RuntimeError("plastic scatter failed")
    return eq, isv
