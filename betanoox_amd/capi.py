"""ctypes binding of libbeta_hip.so (include/beta_hip.h).

The product path: every call goes to the HIP library.  If the library is
missing or no GPU is visible the constructor raises; there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libbeta_hip.so")

# status codes (beta_hip.h)
OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_OOM = -3
ERR_NOT_CONVERGED = -4
ERR_ZERO_DIAG = -5
ERR_NEG_DIAG = -6
ERR_NEG_DETJ = -7
ERR_STATE = -8
ERR_UNSUPPORTED = -9
ERR_CAPACITY = -10
TANGENT_CONTINUUM = 0
TANGENT_ALGORITHMIC = 1

MAT_ISO_ELASTIC = 1
MAT_CMAT3D = 2
MAT_PLANE_STRESS = 3
MAT_HEAT = 4
MAT_J2 = 5

ASM_STIFFNESS = 1
ASM_HEAT_TRANSIENT = 2
ASM_PLASTIC_TANGENT = 3

STOP_REF_R_OVER_X = 0
STOP_REL_RESID = 1
PREC_FP64 = 0
PREC_FP32_IR = 1
PREC_FP64_BLOCKED = 2

FMT_RCV_LOWER = 1
FMT_RCV_UPPER = 2
FMT_CSR_ASCII = 3
FMT_CSR_BINARY = 4
FMT_RCV = 5
TRI_LOWER = 1
TRI_UPPER = 2
TRI_FULL = 3
VEC_LOAD = 1
VEC_SOLUTION = 2

HEAT_X0_ZERO = 0
HEAT_X0_PREVIOUS = 1
HEAT_X0_REFERENCE = 2
HEAT_X0_PROJECTED = 3

# every symbol the header declares
EXPORTS = [
    "bhip_create", "bhip_destroy", "bhip_last_error", "bhip_version", "bhip_set_mesh",
    "bhip_num_dof", "bhip_get_dof_first", "bhip_set_restraints", "bhip_set_prescribed",
    "bhip_build_pattern", "bhip_pattern_nnz", "bhip_get_pattern_lower", "bhip_get_values_lower",
    "bhip_get_csr", "bhip_set_materials", "bhip_set_integration", "bhip_assemble",
    "bhip_get_pseudo_load", "bhip_set_load", "bhip_build_rhs", "bhip_get_rhs", "bhip_set_rhs",
    "bhip_set_x0", "bhip_solve", "bhip_get_solution", "bhip_pcg_solve", "bhip_pcg_solve_noalloc", "bhip_spmv",
    "bhip_internal_force", "bhip_set_zero_diag_replacement",
    "bhip_solve_cases",
    "bhip_bench_spmv", "bhip_bench_stream", "bhip_bench_read", "bhip_plastic_init", "bhip_plastic_update", "bhip_plastic_residual",
    "bhip_plastic_commit", "bhip_plastic_increment", "bhip_plastic_get_state", "bhip_plastic_set_tangent",
    "bhip_server_create", "bhip_server_destroy", "bhip_server_submit", "bhip_server_poll", "bhip_server_wait",
    "bhip_server_budget", "bhip_request_footprint", "bhip_heat_run", "bhip_write_matrix", "bhip_write_csr_file",
    "bhip_set_csr", "bhip_read_matrix", "bhip_write_vector", "bhip_read_vector",
    "bhip_tune_spmv", "bhip_set_spmv_shape", "bhip_get_spmv_shape", "bhip_spmv_shape_describe",
    "bhip_pack_create", "bhip_pack_destroy", "bhip_pack_solve_cases",
]


class SolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxit", C.c_int), ("stop_rule", C.c_int),
                ("precision", C.c_int), ("inner_tol", C.c_double), ("max_outer", C.c_int),
                ("check_every", C.c_int), ("use_x0", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int), ("outer_iters", C.c_int), ("ref_norm", C.c_double),
                ("rel_resid", C.c_double), ("ms_total", C.c_double), ("ms_spmv_avg", C.c_double),
                ("spmv_launches", C.c_int), ("spmv_bytes", C.c_double), ("case_iters", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerOpts(C.Structure):
    _fields_ = [("max_slots", C.c_int), ("hbm_budget", C.c_double), ("tol", C.c_double),
                ("inner_tol", C.c_double), ("maxit", C.c_int), ("precision", C.c_int), ("max_pack", C.c_int)]


class Request(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("xyz", C.POINTER(C.c_double)), ("n_el", C.c_int),
                ("conn", C.POINTER(C.c_int)), ("rest", C.POINTER(C.c_ubyte)), ("mat_kind", C.c_int),
                ("n_groups", C.c_int), ("mat", C.POINTER(C.c_double)), ("elem_group", C.POINTER(C.c_int)),
                ("n_cases", C.c_int), ("presc", C.POINTER(C.c_double)), ("load", C.POINTER(C.c_double)),
                ("order", C.c_int)]


class RequestStats(C.Structure):
    _fields_ = [("status", C.c_int), ("slot", C.c_int), ("reused_pattern", C.c_int), ("iters", C.c_int),
                ("outer_iters", C.c_int), ("rel_resid", C.c_double), ("queue_ms", C.c_double),
                ("service_ms", C.c_double), ("latency_ms", C.c_double), ("start_ms", C.c_double),
                ("end_ms", C.c_double), ("pack_size", C.c_int), ("setup_ms", C.c_double),
                ("solve_ms", C.c_double), ("case_iters", C.c_longlong), ("message", C.c_char * 256)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "message"}
        d["message"] = self.message.decode(errors="replace")
        return d


class HeatStats(C.Structure):
    _fields_ = [("steps", C.c_int), ("iters", C.c_int), ("outer_iters", C.c_int), ("max_rel_resid", C.c_double),
                ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PlasticStats(C.Structure):
    _fields_ = [("newton_iters", C.c_int), ("pcg_iters", C.c_int), ("max_residual", C.c_double),
                ("plastic_points", C.c_int), ("max_return_iters", C.c_int), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SpmvShape(C.Structure):
    _fields_ = [("id", C.c_int), ("slices_per_block", C.c_int), ("unroll", C.c_int), ("nontemporal", C.c_int),
                ("column_bytes", C.c_int), ("ms", C.c_double), ("gbps", C.c_double), ("iter_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BhipError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("status %d: %s" % (status, msg))
        self.status = status


_lib = None

_P_D = C.POINTER(C.c_double)
_P_I = C.POINTER(C.c_int)
_P_U8 = C.POINTER(C.c_ubyte)
_P_LL = C.POINTER(C.c_longlong)


def load_library(path=LIB_PATH):
    """Load the HIP library and declare its signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError("libbeta_hip.so not built at %s (run __graft_entry__.build())" % path)
    lib = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "bhip_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "bhip_destroy": (None, [vp]),
        "bhip_last_error": (C.c_char_p, [vp]),
        "bhip_version": (C.c_char_p, []),
        "bhip_set_mesh": (C.c_int, [vp, C.c_int, C.c_int, _P_D, C.c_int, C.c_int, _P_I, _P_I, C.c_int]),
        "bhip_num_dof": (C.c_int, [vp]),
        "bhip_get_dof_first": (C.c_int, [vp, _P_I]),
        "bhip_set_restraints": (C.c_int, [vp, _P_U8, _P_D]),
        "bhip_set_prescribed": (C.c_int, [vp, _P_D]),
        "bhip_build_pattern": (C.c_int, [vp]),
        "bhip_pattern_nnz": (C.c_int, [vp, _P_LL, _P_LL]),
        "bhip_get_pattern_lower": (C.c_int, [vp, _P_I, _P_I]),
        "bhip_get_values_lower": (C.c_int, [vp, _P_D]),
        "bhip_get_csr": (C.c_int, [vp, _P_I, _P_I, _P_D]),
        "bhip_set_materials": (C.c_int, [vp, C.c_int, C.c_int, _P_D, _P_I]),
        "bhip_set_integration": (C.c_int, [vp, C.c_int, C.c_double]),
        "bhip_assemble": (C.c_int, [vp, C.c_int, _P_D, _P_D]),
        "bhip_get_pseudo_load": (C.c_int, [vp, _P_D]),
        "bhip_set_load": (C.c_int, [vp, _P_D]),
        "bhip_build_rhs": (C.c_int, [vp]),
        "bhip_get_rhs": (C.c_int, [vp, _P_D]),
        "bhip_set_rhs": (C.c_int, [vp, _P_D]),
        "bhip_set_x0": (C.c_int, [vp, _P_D]),
        "bhip_solve": (C.c_int, [vp, C.POINTER(SolveOpts), C.POINTER(Stats)]),
        "bhip_get_solution": (C.c_int, [vp, _P_D]),
        "bhip_pcg_solve": (C.c_int, [vp, _P_D, _P_D, C.POINTER(SolveOpts), C.POINTER(Stats)]),
        "bhip_pcg_solve_noalloc": (C.c_int, [vp, _P_D, _P_D, C.POINTER(SolveOpts), C.POINTER(Stats)]),
        "bhip_spmv": (C.c_int, [vp, _P_D, _P_D]),
        "bhip_internal_force": (C.c_int, [vp, C.c_int, _P_D, _P_D, _P_D, _P_D]),
        "bhip_set_zero_diag_replacement": (C.c_int, [vp, C.c_double]),
        "bhip_solve_cases": (C.c_int, [vp, C.c_int, _P_D, _P_D, _P_D, C.POINTER(SolveOpts), C.POINTER(Stats)]),
        "bhip_bench_spmv": (C.c_int, [vp, C.c_int, C.c_int, _P_D, _P_D]),
        "bhip_bench_stream": (C.c_int, [vp, C.c_longlong, C.c_int, _P_D, _P_D]),
        "bhip_bench_read": (C.c_int, [vp, C.c_longlong, C.c_int, _P_D, _P_D]),
        "bhip_tune_spmv": (C.c_int, [vp, C.c_int, C.POINTER(SpmvShape), C.c_int, _P_I, _P_I]),
        "bhip_set_spmv_shape": (C.c_int, [vp, C.c_int]),
        "bhip_get_spmv_shape": (C.c_int, [vp]),
        "bhip_spmv_shape_describe": (C.c_int, [C.c_int, C.POINTER(SpmvShape)]),
        "bhip_pack_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "bhip_pack_destroy": (None, [vp]),
        "bhip_pack_solve_cases": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.c_int, C.POINTER(_P_D), C.POINTER(_P_D),
                                            C.POINTER(_P_D), C.POINTER(SolveOpts), C.POINTER(Stats), _P_I]),
        "bhip_plastic_init": (C.c_int, [vp]),
        "bhip_plastic_set_tangent": (C.c_int, [vp, C.c_int]),
        "bhip_plastic_update": (C.c_int, [vp, _P_D]),
        "bhip_plastic_residual": (C.c_int, [vp, _P_D, _P_D]),
        "bhip_plastic_commit": (C.c_int, [vp]),
        "bhip_plastic_increment": (C.c_int, [vp, _P_D, C.c_double, C.c_int, C.POINTER(SolveOpts),
                                             C.POINTER(PlasticStats)]),
        "bhip_plastic_get_state": (C.c_int, [vp, _P_D, _P_D, _P_D]),
        "bhip_server_create": (C.c_int, [C.c_int, C.POINTER(ServerOpts), C.POINTER(vp)]),
        "bhip_server_destroy": (None, [vp]),
        "bhip_server_submit": (C.c_int, [vp, C.POINTER(Request), _P_LL]),
        "bhip_server_poll": (C.c_int, [vp, C.c_longlong, _P_I]),
        "bhip_server_wait": (C.c_int, [vp, C.c_longlong, _P_D, C.POINTER(RequestStats)]),
        "bhip_server_budget": (C.c_double, [vp]),
        "bhip_request_footprint": (C.c_double, [C.c_int, C.c_int, C.c_int]),
        "bhip_heat_run": (C.c_int, [vp, C.c_int, _P_D, C.c_int, C.POINTER(SolveOpts), C.POINTER(HeatStats)]),
        "bhip_write_matrix": (C.c_int, [vp, C.c_char_p, C.c_int]),
        "bhip_write_csr_file": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_longlong, _P_I, _P_I, _P_D]),
        "bhip_set_csr": (C.c_int, [vp, C.c_int, C.c_longlong, _P_I, _P_I, _P_D, C.c_int, C.c_int]),
        "bhip_read_matrix": (C.c_int, [vp, C.c_char_p, C.c_int]),
        "bhip_write_vector": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _P_D]),
        "bhip_read_vector": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _P_D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_P_D)


def _i(a):
    return None if a is None else a.ctypes.data_as(_P_I)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    """One device context: mesh, pattern, matrix, solver state on one GPU."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.bhip_create(device, C.byref(h))
        if st != OK:
            raise BhipError(st, "bhip_create failed (is a GPU visible?)")
        self.h = h
        self.n_dof = 0

    def close(self):
        if self.h:
            self.lib.bhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, st, allow=()):
        if st != OK and st not in allow:
            raise BhipError(st, self.lib.bhip_last_error(self.h).decode())
        return st

    # ---- mesh / restraints
    def set_mesh(self, dim, xyz, conn, ndof_per_node=None, ndof_uniform=None):
        xyz = _f64(xyz)
        conn = _i32(conn)
        nn = xyz.shape[0]
        ne, nnpe = conn.shape
        nd = None if ndof_per_node is None else _i32(ndof_per_node)
        if ndof_uniform is None:
            ndof_uniform = dim
        self._keep = (xyz, conn, nd)
        self._chk(self.lib.bhip_set_mesh(self.h, dim, nn, _d(xyz), ne, nnpe, _i(conn), _i(nd), ndof_uniform))
        self.n_dof = self.lib.bhip_num_dof(self.h)
        return self.n_dof

    def dof_first(self):
        out = np.zeros(len(self._keep[0]), dtype=np.int32)
        self._chk(self.lib.bhip_get_dof_first(self.h, _i(out)))
        return out

    def set_restraints(self, restrained, values=None):
        r = np.ascontiguousarray(restrained, dtype=np.uint8)
        v = None if values is None else _f64(values)
        self._chk(self.lib.bhip_set_restraints(self.h, r.ctypes.data_as(_P_U8), _d(v)))

    def set_prescribed(self, values):
        self._chk(self.lib.bhip_set_prescribed(self.h, _d(_f64(values))))

    # ---- pattern
    def build_pattern(self):
        self._chk(self.lib.bhip_build_pattern(self.h))

    def nnz(self):
        lo, full = C.c_longlong(), C.c_longlong()
        self._chk(self.lib.bhip_pattern_nnz(self.h, C.byref(lo), C.byref(full)))
        return lo.value, full.value

    def pattern_lower(self):
        nl, _ = self.nnz()
        rp = np.zeros(self.n_dof + 1, dtype=np.int32)
        ci = np.zeros(max(nl, 1), dtype=np.int32)
        self._chk(self.lib.bhip_get_pattern_lower(self.h, _i(rp), _i(ci)))
        return rp, ci[:nl]

    def values_lower(self):
        nl, _ = self.nnz()
        v = np.zeros(max(nl, 1))
        self._chk(self.lib.bhip_get_values_lower(self.h, _d(v)))
        return v[:nl]

    def csr(self):
        _, nf = self.nnz()
        rp = np.zeros(self.n_dof + 1, dtype=np.int32)
        ci = np.zeros(max(nf, 1), dtype=np.int32)
        v = np.zeros(max(nf, 1))
        self._chk(self.lib.bhip_get_csr(self.h, _i(rp), _i(ci), _d(v)))
        return rp, ci[:nf], v[:nf]

    # ---- materials / assembly
    def set_materials(self, kind, params, elem_group=None):
        p = _f64(params)
        ng = p.shape[0] if p.ndim == 2 else 1
        g = None if elem_group is None else _i32(elem_group)
        self._chk(self.lib.bhip_set_materials(self.h, kind, ng, _d(p), _i(g)))

    def set_integration(self, order, factor=1.0):
        self._chk(self.lib.bhip_set_integration(self.h, order, factor))

    def assemble(self, what=ASM_STIFFNESS, coef=None, u_state=None):
        c = None if coef is None else _f64(coef)
        u = None if u_state is None else _f64(u_state)
        self._chk(self.lib.bhip_assemble(self.h, what, _d(c), _d(u)))

    def pseudo_load(self):
        f = np.zeros(self.n_dof)
        self._chk(self.lib.bhip_get_pseudo_load(self.h, _d(f)))
        return f

    def set_load(self, load):
        self._chk(self.lib.bhip_set_load(self.h, _d(None if load is None else _f64(load))))

    def build_rhs(self):
        self._chk(self.lib.bhip_build_rhs(self.h))

    def rhs(self):
        b = np.zeros(self.n_dof)
        self._chk(self.lib.bhip_get_rhs(self.h, _d(b)))
        return b

    def set_rhs(self, b):
        self._chk(self.lib.bhip_set_rhs(self.h, _d(_f64(b))))

    def set_x0(self, x0):
        self._chk(self.lib.bhip_set_x0(self.h, _d(None if x0 is None else _f64(x0))))

    # ---- solve
    @staticmethod
    def opts(tol=1e-8, maxit=10000, stop_rule=STOP_REL_RESID, precision=PREC_FP64, inner_tol=1e-3,
             max_outer=50, check_every=32, use_x0=0):
        return SolveOpts(tol, maxit, stop_rule, precision, inner_tol, max_outer, check_every, use_x0)

    def solve(self, opts=None, allow_not_converged=False, **kw):
        o = opts if opts is not None else self.opts(**kw)
        s = Stats()
        allow = (ERR_NOT_CONVERGED,) if allow_not_converged else ()
        st = self._chk(self.lib.bhip_solve(self.h, C.byref(o), C.byref(s)), allow)
        d = s.as_dict()
        d["status"] = st
        return d

    def solution(self):
        x = np.zeros(self.n_dof)
        self._chk(self.lib.bhip_get_solution(self.h, _d(x)))
        return x

    def pcg_solve(self, b, x0=None, opts=None, **kw):
        """LargeMatrix::solve(b, x0) mirror: returns (solution, stats)."""
        bb = _f64(b).copy()
        o = opts if opts is not None else self.opts(**kw)
        s = Stats()
        st = self._chk(self.lib.bhip_pcg_solve(self.h, _d(bb), _d(None if x0 is None else _f64(x0)),
                                               C.byref(o), C.byref(s)), (ERR_NOT_CONVERGED,))
        d = s.as_dict()
        d["status"] = st
        return bb, d

    def internal_force(self, u):
        """BasicModel::getGlobalForces: (force, residual, max |residual|) for
        one displacement vector or each row of an (nvec, n_dof) array."""
        U = np.ascontiguousarray(np.atleast_2d(u), dtype=np.float64)
        if U.shape[1] != self.n_dof:
            raise ValueError("u must be (nvec, n_dof)")
        F, R, M = np.zeros_like(U), np.zeros_like(U), np.zeros(U.shape[0])
        self._chk(self.lib.bhip_internal_force(self.h, U.shape[0], _d(U), _d(F), _d(R), _d(M)))
        if np.ndim(u) == 1:
            return F[0], R[0], M[0]
        return F, R, M

    def set_zero_diag_replacement(self, value):
        """LargeMatrix::setReplaceZeroDiagonal_Option (0 = report zero diagonals)."""
        self._chk(self.lib.bhip_set_zero_diag_replacement(self.h, value))

    def solve_cases(self, presc, load=None, opts=None, allow_not_converged=False, **kw):
        """Solve every load case of `presc` ((nrhs, n_dof) prescribed values)
        with the assembled stiffness; returns (X (nrhs, n_dof), stats)."""
        P = np.ascontiguousarray(np.atleast_2d(presc), dtype=np.float64)
        if P.shape[1] != self.n_dof:
            raise ValueError("presc must be (nrhs, n_dof)")
        X = np.zeros_like(P)
        o = opts if opts is not None else self.opts(**kw)
        s = Stats()
        allow = (ERR_NOT_CONVERGED,) if allow_not_converged else ()
        st = self._chk(self.lib.bhip_solve_cases(self.h, P.shape[0], _d(P), _d(None if load is None else _f64(load)),
                                                 _d(X), C.byref(o), C.byref(s)), allow)
        d = s.as_dict()
        d["status"] = st
        return X, d

    def heat_run(self, n_steps, u0, x0_mode=1, opts=None, **kw):
        """bhip_heat_run: n_steps of the transient diffusion after a
        BHIP_ASM_HEAT_TRANSIENT assembly; returns (u, stats)."""
        u = _f64(u0).copy()
        o = opts if opts is not None else self.opts(**kw)
        s = HeatStats()
        self._chk(self.lib.bhip_heat_run(self.h, n_steps, _d(u), x0_mode, C.byref(o), C.byref(s)))
        return u, s.as_dict()

    # ---- exchange formats
    def write_matrix(self, path, fmt):
        self._chk(self.lib.bhip_write_matrix(self.h, str(path).encode(), fmt))

    def set_csr(self, ia, ja, a, base=0, triangle=TRI_FULL):
        ia, ja, a = _i32(ia), _i32(ja), _f64(a)
        n = ia.shape[0] - 1
        self._chk(self.lib.bhip_set_csr(self.h, n, ja.shape[0], ia.ctypes.data_as(_P_I), ja.ctypes.data_as(_P_I),
                                        _d(a), base, triangle))
        self.n_dof = n

    def read_matrix(self, path, fmt):
        self._chk(self.lib.bhip_read_matrix(self.h, str(path).encode(), fmt))
        self.n_dof = self.lib.bhip_num_dof(self.h)

    def spmv(self, x):
        y = np.zeros(self.n_dof)
        self._chk(self.lib.bhip_spmv(self.h, _d(_f64(x)), _d(y)))
        return y

    # ---- plasticity
    def plastic_init(self):
        self._chk(self.lib.bhip_plastic_init(self.h))

    def plastic_set_tangent(self, tangent):
        """TANGENT_CONTINUUM (the reference's C_ep) or TANGENT_ALGORITHMIC."""
        self._chk(self.lib.bhip_plastic_set_tangent(self.h, tangent))

    def plastic_update(self, du):
        self._chk(self.lib.bhip_plastic_update(self.h, _d(_f64(du))))

    def plastic_residual(self):
        r = np.zeros(self.n_dof)
        m = C.c_double()
        self._chk(self.lib.bhip_plastic_residual(self.h, _d(r), C.byref(m)))
        return r, m.value

    def plastic_commit(self):
        self._chk(self.lib.bhip_plastic_commit(self.h))

    def plastic_increment(self, presc_increment, max_allowable_residual, max_iterations, opts=None, **kw):
        o = opts if opts is not None else self.opts(**kw)
        s = PlasticStats()
        p = None if presc_increment is None else _f64(presc_increment)
        st = self._chk(self.lib.bhip_plastic_increment(self.h, _d(p), max_allowable_residual, max_iterations,
                                                        C.byref(o), C.byref(s)), (ERR_NOT_CONVERGED,))
        d = s.as_dict()
        d["status"] = st
        return d

    def plastic_state(self, n_el, nip):
        s = np.zeros(n_el * nip * 6)
        e = np.zeros(n_el * nip)
        u = np.zeros(self.n_dof)
        self._chk(self.lib.bhip_plastic_get_state(self.h, _d(s), _d(e), _d(u)))
        return s.reshape(n_el, nip, 6), e.reshape(n_el, nip), u

    def bench_stream(self, nbytes=1 << 30, reps=20):
        ms, g = C.c_double(), C.c_double()
        self._chk(self.lib.bhip_bench_stream(self.h, nbytes, reps, C.byref(ms), C.byref(g)))
        return ms.value, g.value

    def bench_read(self, nbytes=2 << 30, reps=10):
        """bhip_bench_read: mean ms and GB/s of a read-only device stream."""
        ms, g = C.c_double(), C.c_double()
        self._chk(self.lib.bhip_bench_read(self.h, nbytes, reps, C.byref(ms), C.byref(g)))
        return ms.value, g.value

    def bench_spmv(self, fp32=False, reps=20, blocked64=False):
        """Mean ms and algorithmic bytes of one SpMV: FP64 CSR, the FP32 form of
        the inner solve (fp32=True) or FP64 SELL blocks (blocked64=True)."""
        ms, by = C.c_double(), C.c_double()
        form = 2 if blocked64 else int(bool(fp32))
        self._chk(self.lib.bhip_bench_spmv(self.h, form, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def tune_spmv(self, reps=10):
        """bhip_tune_spmv: time every SELL tile shape on the assembled matrix,
        keep the fastest.  Returns (chosen id or -1, list of shape dicts)."""
        table = (SpmvShape * 64)()
        n, chosen = C.c_int(), C.c_int()
        self._chk(self.lib.bhip_tune_spmv(self.h, reps, table, 64, C.byref(n), C.byref(chosen)))
        return chosen.value, [table[i].as_dict() for i in range(n.value)]

    def set_spmv_shape(self, shape_id):
        self._chk(self.lib.bhip_set_spmv_shape(self.h, int(shape_id)))

    def spmv_shape(self):
        return self.lib.bhip_get_spmv_shape(self.h)


class Server:
    """Resident request loop (bhip_server_*): hex8 elasticity requests with
    several prescribed-displacement load cases, solved by up to `max_slots`
    workers at once on one GPU; each worker packs up to `max_pack` small
    requests into one batched solve."""

    def __init__(self, device=0, max_slots=4, hbm_budget=0.0, tol=1e-8, inner_tol=1e-3, maxit=40000,
                 precision=PREC_FP32_IR, max_pack=0):
        self.lib = load_library()
        o = ServerOpts(max_slots, hbm_budget, tol, inner_tol, maxit, precision, max_pack)
        h = C.c_void_p()
        st = self.lib.bhip_server_create(device, C.byref(o), C.byref(h))
        if st != OK:
            raise BhipError(st, "bhip_server_create failed (is a GPU visible?)")
        self.h = h
        self._shape = {}

    def close(self):
        if self.h:
            self.lib.bhip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def budget(self):
        return self.lib.bhip_server_budget(self.h)

    def submit(self, prob, cases=None, elem_group=None):
        """Queue a hex8 Problem; `cases` ((n_cases, n_dof) prescribed values)
        defaults to prob.extra["cases"], elem_group to prob.extra["group"]."""
        P = np.ascontiguousarray(np.atleast_2d(cases if cases is not None else np.stack(prob.extra["cases"])),
                                 dtype=np.float64)
        xyz, conn = _f64(prob.xyz), _i32(prob.conn)
        rest = np.ascontiguousarray(prob.restrained, dtype=np.uint8)
        mat = _f64(prob.mat)
        g = elem_group if elem_group is not None else prob.extra.get("group")
        g = None if g is None else _i32(g)
        load = None if prob.load is None or not np.any(prob.load) else _f64(prob.load)
        q = Request(prob.n_nodes, _d(xyz), prob.n_el, conn.ctypes.data_as(_P_I), rest.ctypes.data_as(_P_U8),
                    prob.mat_kind, mat.shape[0], _d(mat), None if g is None else g.ctypes.data_as(_P_I),
                    P.shape[0], _d(P), _d(load), prob.order)
        rid = C.c_longlong()
        st = self.lib.bhip_server_submit(self.h, C.byref(q), C.byref(rid))
        if st != OK:
            raise BhipError(st, "bhip_server_submit failed")
        self._shape[rid.value] = P.shape
        return rid.value

    def poll(self, rid):
        d = C.c_int()
        st = self.lib.bhip_server_poll(self.h, rid, C.byref(d))
        if st != OK:
            raise BhipError(st, "unknown request id")
        return bool(d.value)

    def wait(self, rid):
        """(X (n_cases, n_dof), stats dict); raises if the request failed."""
        shape = self._shape.pop(rid)
        X = np.zeros(shape)
        s = RequestStats()
        st = self.lib.bhip_server_wait(self.h, rid, _d(X), C.byref(s))
        d = s.as_dict()
        if st != OK:
            raise BhipError(st, d["message"])
        return X, d


class Pack:
    """Packed solver (bhip_pack_*): the load cases of several assembled
    Contexts solved together in one batched FP32-inner PCG."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.bhip_pack_create(device, C.byref(h))
        if st != OK:
            raise BhipError(st, "bhip_pack_create failed (is a GPU visible?)")
        self.h = h

    def close(self):
        if self.h:
            self.lib.bhip_pack_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def solve_cases(self, ctxs, prescs, loads=None, opts=None, **kw):
        """ctxs: list of Context; prescs: list of (nrhs, n_dof) arrays (same
        nrhs).  Returns (list of X, list of stats dicts, list of status)."""
        o = opts if opts is not None else Context.opts(**kw)
        n = len(ctxs)
        P = [np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64) for a in prescs]
        nrhs = P[0].shape[0]
        X = [np.zeros_like(a) for a in P]
        L = None
        if loads is not None:
            L = [None if l is None else _f64(l) for l in loads]
        handles = (C.c_void_p * n)(*[c.h for c in ctxs])
        pp = (_P_D * n)(*[_d(a) for a in P])
        xp = (_P_D * n)(*[_d(a) for a in X])
        lp = None if L is None else (_P_D * n)(*[_d(l) for l in L])
        stats = (Stats * n)()
        status = np.zeros(n, dtype=np.int32)
        self.lib.bhip_pack_solve_cases(self.h, n, handles, nrhs, pp, lp, xp, C.byref(o), stats,
                                       status.ctypes.data_as(_P_I))
        return X, [s.as_dict() for s in stats], [int(v) for v in status]


def spmv_shapes():
    """Every SELL tile shape the library knows (bhip_spmv_shape_describe)."""
    lib = load_library()
    out = []
    info = SpmvShape()
    while lib.bhip_spmv_shape_describe(len(out), C.byref(info)) == 0:
        out.append(info.as_dict())
    return out


def write_csr_file(path, fmt, ia, ja, a):
    """bhip_write_csr_file: raw arrays in a reference file format (no GPU)."""
    lib = load_library()
    ia, ja, a = _i32(ia), _i32(ja), _f64(a)
    st = lib.bhip_write_csr_file(str(path).encode(), fmt, ia.shape[0] - 1, ja.shape[0], ia.ctypes.data_as(_P_I),
                                 ja.ctypes.data_as(_P_I), _d(a))
    if st != OK:
        raise BhipError(st, "bhip_write_csr_file failed")


def write_vector(path, kind, v):
    lib = load_library()
    v = _f64(v)
    st = lib.bhip_write_vector(str(path).encode(), kind, v.shape[0], _d(v))
    if st != OK:
        raise BhipError(st, "bhip_write_vector failed")


def read_vector(path, kind, n):
    lib = load_library()
    v = np.zeros(n)
    st = lib.bhip_read_vector(str(path).encode(), kind, n, _d(v))
    if st != OK:
        raise BhipError(st, "bhip_read_vector failed")
    return v
