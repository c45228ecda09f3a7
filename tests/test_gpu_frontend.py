"""A BETA mesh file and solver-settings commands run through the C++
front-end (beta_hip_driver mesh ...) give the same analysis as the built-in
structured meshes of the driver, and the oracle's solution.

Tolerances: the SPARSE settings (FP64 PCG, reference stop rule) reproduce the
structured run's iteration count exactly and its solution to 1e-12 (the
nodal loads are summed face by face); HIPPCG (FP32 inner + FP64 refinement)
meets its 1e-8 relative residual and agrees with the oracle to 1e-6, and
HIPPCG_FP64_BLOCKED meets 1e-11 and agrees to 1e-8; the
Solution.txt the front-end writes reads back to the --dump of the same run.
"""
import os
import subprocess

import numpy as np
import pytest

from betanoox_amd import capi, meshgen, meshio
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "betanoox_amd", "_build", "beta_hip_driver")


def _kv(args):
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return dict(line.split("=", 1) for line in r.stdout.split() if "=" in line)


@pytest.mark.parametrize("kind,n,binary", [("hex", 5, False), ("hex", 4, True), ("quad", 12, False)])
def test_mesh_file_matches_structured_run(tmp_path, kind, n, binary):
    print("case", kind, n, binary, flush=True)
    if kind == "hex":
        xyz, conn = meshgen.structured_hex8(n, n, n)
    else:
        xyz, conn = meshgen.structured_quad4(n, n)
    mfile = tmp_path / ("m.bin" if binary else "m.txt")
    (meshio.write_binary if binary else meshio.write_text)(mfile, xyz, conn, base=1)
    settings = tmp_path / "s.txt"
    settings.write_text("setStorageMethod SPARSE\nsetSparseSolverTolerance 1e-9\nsetSparseSolverMaxIterations 20000\nend\n")
    args = ["mesh", mfile, "--settings", settings, "--solution", tmp_path / "Solution.txt"]
    if binary:
        args.append("--binary")
    got = _kv(args)
    ref = _kv([kind, n, "--tol", "1e-9", "--rule", "ref", "--dump", tmp_path / "u.bin"])
    assert got["n_dof"] == ref["n_dof"] and got["iters"] == ref["iters"]
    assert abs(float(got["tip"]) - float(ref["tip"])) <= 1e-12 * abs(float(ref["tip"]))
    u_ref = np.fromfile(tmp_path / "u.bin", dtype=np.float64)
    u = capi.read_vector(tmp_path / "Solution.txt", capi.VEC_SOLUTION, u_ref.shape[0])
    assert np.abs(u - u_ref).max() <= 1e-12 * np.abs(u_ref).max()


@pytest.mark.parametrize("token,tol,err", [("HIPPCG", 1e-8, 1e-6), ("HIPPCG_FP64_BLOCKED", 1e-11, 1e-8)])
def test_hippcg_settings_against_oracle(tmp_path, token, tol, err):
    """HIPPCG (FP32 inner + FP64 refinement) to 1e-8 agrees with the oracle's
    1e-12 solve to 1e-6; HIPPCG_FP64_BLOCKED (FP64 PCG on FP64 node blocks)
    to 1e-11 agrees to 1e-8."""
    n = 5
    xyz, conn = meshgen.structured_hex8(n, n, n)
    meshio.write_text(tmp_path / "m.txt", xyz, conn)
    (tmp_path / "s.txt").write_text("setStorageMethod %s\nsetSparseSolverTolerance %g\nend\n" % (token, tol))
    got = _kv(["mesh", tmp_path / "m.txt", "--settings", tmp_path / "s.txt", "--solution", tmp_path / "sol.txt"])
    assert float(got["rel_resid"]) <= tol and int(got["outer"]) >= 1
    prob = meshgen.elastic_cube(n, jitter=0.0, e_range=(1.0, 1.0), nu_range=(0.3, 0.3))
    eq = O.assemble_problem(prob)
    st, x, it, _ = eq.pcg(eq.rhs(), eps=1e-12, stop_rule=1)
    u = capi.read_vector(tmp_path / "sol.txt", capi.VEC_SOLUTION, prob.n_dof)
    assert np.linalg.norm(u - x) <= err * np.linalg.norm(x)
