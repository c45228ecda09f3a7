"""Mesh files and solver-settings commands (no GPU): the Python mesh module
and the C++ front-end (beta_hip_driver mesh ... --parse-only) read the
BasicMesh::ReadMesh / ReadMeshBinary layouts (BasicMesh.cpp:79-268) and the
Equations::processCommand subset (equation.cpp:1769-1903)."""
import os
import subprocess

import numpy as np
import pytest

from betanoox_amd import meshgen, meshio

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "betanoox_amd", "_build",
                      "beta_hip_driver")


def _parse(*args):
    r = subprocess.run([DRIVER, "mesh"] + [str(a) for a in args] + ["--parse-only"], capture_output=True, text=True,
                       timeout=60)
    out = dict(l.split("=", 1) for l in r.stdout.split() if "=" in l)
    return r.returncode, out, r.stderr


@pytest.mark.parametrize("base", [0, 1])
def test_python_round_trips(tmp_path, base):
    xyz, conn = meshgen.structured_hex8(3, 2, 2)
    xyz = xyz + 0.01 * np.sin(np.arange(xyz.size)).reshape(xyz.shape)
    meshio.write_text(tmp_path / "m.txt", xyz, conn, base=base)
    x2, c2, b2 = meshio.read_text(tmp_path / "m.txt")
    assert b2 == base and np.array_equal(c2, conn) and np.array_equal(x2, xyz)
    meshio.write_binary(tmp_path / "m.bin", xyz, conn, base=base)
    x3, c3, b3 = meshio.read_binary(tmp_path / "m.bin")
    assert b3 == base and np.array_equal(c3, conn) and np.array_equal(x3, xyz)
    # 2D mesh written with an explicit read option block
    xy, q = meshgen.structured_quad4(3, 2)
    meshio.write_text(tmp_path / "q.txt", xy, q, base=base, read_option="xy")
    x4, c4, _ = meshio.read_text(tmp_path / "q.txt")
    assert np.array_equal(x4, xy) and np.array_equal(c4, q)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="beta_hip_driver not built")
def test_cpp_reader_and_settings(tmp_path):
    xyz, conn = meshgen.structured_hex8(4, 3, 2)
    meshio.write_text(tmp_path / "m.txt", xyz, conn, base=1)
    meshio.write_binary(tmp_path / "m.bin", xyz, conn, base=0)
    (tmp_path / "s.txt").write_text("setStorageMethod HIPPCG\nsetSparseSolverTolerance 1e-9\n"
                                    "setSparseSolverMaxIterations 500\nUseMultiCore max\nend\n")
    rc, out, err = _parse(tmp_path / "m.txt", "--settings", tmp_path / "s.txt")
    assert rc == 0, err
    assert out["n_nodes"] == str(xyz.shape[0]) and out["n_el"] == str(conn.shape[0])
    assert out["num_dims"] == "3" and out["node_offset"] == "1"
    assert out["storage"] == "HIPPCG" and out["precision"] == "1" and out["max_iterations"] == "500"
    assert float(out["tolerance"]) == 1e-9 and out["parallel_assembly_hip"] == "1"
    rc, out, err = _parse(tmp_path / "m.bin", "--binary", "--parallel-assembly", "ESA")
    assert rc == 0, err
    assert out["node_offset"] == "0" and out["storage"] == "SPARSE" and out["precision"] == "0"
    assert out["parallel_assembly_hip"] == "0"
    for token, prec in (("HIPPCG_FP64", "0"), ("HIPPCG_FP64_BLOCKED", "2"), ("hippcg", "1")):
        (tmp_path / "s2.txt").write_text("setStorageMethod %s\nend\n" % token)
        rc, out, err = _parse(tmp_path / "m.bin", "--binary", "--settings", tmp_path / "s2.txt")
        assert rc == 0, err
        assert out["storage"] == token.upper() and out["precision"] == prec


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="beta_hip_driver not built")
def test_cpp_reader_rejects_bad_input(tmp_path):
    xyz, conn = meshgen.structured_hex8(2, 2, 2)
    bad = conn.copy()
    bad[3, 2] = 999
    meshio.write_text(tmp_path / "bad.txt", xyz, bad)
    rc, _, err = _parse(tmp_path / "bad.txt")
    assert rc == 1 and "node that does not exist" in err
    meshio.write_text(tmp_path / "ok.txt", xyz, conn)
    (tmp_path / "s.txt").write_text("setStorageMethod NOSUCH\nend\n")
    rc, _, err = _parse(tmp_path / "ok.txt", "--settings", tmp_path / "s.txt")
    assert rc == 1 and "does not exist for this platform" in err
    (tmp_path / "s2.txt").write_text("setStorageMethod SPARSE\nfrobnicate 3\nend\n")
    rc, _, err = _parse(tmp_path / "ok.txt", "--settings", tmp_path / "s2.txt")
    assert rc == 1 and "Did not recognize command" in err
    rc, _, err = _parse(tmp_path / "ok.txt", "--parallel-assembly", "XYZ")
    assert rc == 1 and "not recognized" in err
