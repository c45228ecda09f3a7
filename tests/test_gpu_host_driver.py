"""The C++ adapters (HipEquations / HipAssembler / HipLargeMatrix) drive a
full linear analysis through the C ABI; the solution must match the oracle.
Tolerance: relative displacement error <= 1e-8 for FP64 PCG solved to 1e-11."""
import os
import subprocess

import numpy as np
import pytest

from betanoox_amd import meshgen
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "betanoox_amd", "_build", "beta_hip_driver")


def _run(args, tmp_path):
    out = os.path.join(tmp_path, "u.bin")
    r = subprocess.run([EXE] + args + ["--dump", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    kv = dict(line.split("=", 1) for line in r.stdout.split())
    return kv, np.fromfile(out, dtype=np.float64)


@pytest.mark.parametrize("kind,n", [("quad", 16), ("hex", 5)])
def test_driver_matches_oracle(kind, n, tmp_path):
    print("case", kind, n, flush=True)
    kv, u = _run([kind, str(n), "--tol", "1e-11", "--rule", "rel"], str(tmp_path))
    if kind == "quad":
        prob = meshgen.plane_stress_quad(n)
    else:
        prob = meshgen.elastic_cube(n, jitter=0.0, e_range=(1.0, 1.0), nu_range=(0.3, 0.3))
    eq = O.assemble_problem(prob)
    st, x, it, _ = eq.pcg(eq.rhs(), eps=1e-12, stop_rule=1)
    assert st == 0
    assert int(kv["n_dof"]) == prob.n_dof
    assert np.linalg.norm(u - x) <= 1e-8 * np.linalg.norm(x)
    assert float(kv["rel_resid"]) <= 1e-11


def test_driver_fp32_refined(tmp_path):
    kv, u = _run(["hex", "5", "--tol", "1e-8", "--fp32ir"], str(tmp_path))
    assert float(kv["rel_resid"]) <= 1e-8
    assert int(kv["outer"]) >= 1
