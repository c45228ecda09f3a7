"""BETA mesh files (Python side, for tests and tools).

Text, as BasicMesh::ReadMesh reads it (BasicMesh.cpp:79-188):
    numNodes numElements numDims
    [processCommands / setCoordinateReadOption xyz|xy|yz|x / endCommands]
    one line per node:    number coordinates
    one line per element: number nnpe node numbers
Binary, as BasicMesh::ReadMeshBinary reads it (BasicMesh.cpp:190-268):
    int numNodes, numElements, numDims; per node int number + 3 doubles;
    per element int number, int nnpe, nnpe ints.
Node and element numbers start at `base` (0 or 1); connectivity refers to
node numbers.  The C++ reader is betanoox_amd/host/beta_hip_frontend.cpp.
"""
import numpy as np


def write_text(path, xyz, conn, base=0, read_option=None):
    xyz = np.asarray(xyz, dtype=np.float64)
    conn = np.asarray(conn)
    nn, nd = xyz.shape
    with open(path, "w") as f:
        f.write("%d %d %d\n" % (nn, conn.shape[0], nd))
        if read_option is not None:
            f.write("processCommands\nsetCoordinateReadOption %s\nendCommands\n" % read_option)
        for i in range(nn):
            f.write("%d %s\n" % (i + base, " ".join(repr(float(v)) for v in xyz[i])))
        for e in range(conn.shape[0]):
            f.write("%d %d %s\n" % (e + base, conn.shape[1], " ".join(str(int(v) + base) for v in conn[e])))


def write_binary(path, xyz, conn, base=0):
    xyz = np.asarray(xyz, dtype=np.float64)
    conn = np.asarray(conn, dtype=np.int32)
    nn, nd = xyz.shape
    with open(path, "wb") as f:
        np.array([nn, conn.shape[0], nd], np.int32).tofile(f)
        for i in range(nn):
            np.array([i + base], np.int32).tofile(f)
            c = np.zeros(3)
            c[:nd] = xyz[i]
            c.tofile(f)
        for e in range(conn.shape[0]):
            np.array([e + base, conn.shape[1]], np.int32).tofile(f)
            (conn[e] + base).astype(np.int32).tofile(f)


def read_text(path):
    """Returns (xyz (nn, numDims), conn (0-based, n_el x nnpe), base)."""
    with open(path) as f:
        tok = f.read().split()
    nn, ne, nd = int(tok[0]), int(tok[1]), abs(int(tok[2]))
    k = 3
    option = {1: "x", 2: "xy", 3: "xyz"}[nd]
    if tok[k] == "processCommands":
        k += 1
        while tok[k] != "endCommands":
            if tok[k] == "setCoordinateReadOption":
                option = tok[k + 1]
                k += 1
            k += 1
        k += 1
    comps = {"x": [0], "xy": [0, 1], "yz": [1, 2], "xyz": [0, 1, 2]}[option]
    xyz = np.zeros((nn, 3))
    num = np.zeros(nn, int)
    for i in range(nn):
        num[i] = int(tok[k])
        for j, c in enumerate(comps):
            xyz[i, c] = float(tok[k + 1 + j])
        k += 1 + len(comps)
    base = 1 if num[0] == 1 else 0
    rows = []
    for e in range(ne):
        nnpe = int(tok[k + 1])
        rows.append([int(t) - base for t in tok[k + 2:k + 2 + nnpe]])
        k += 2 + nnpe
    return xyz[:, :nd], np.array(rows, np.int32), base


def read_binary(path):
    data = open(path, "rb").read()
    nn, ne, nd = np.frombuffer(data, np.int32, 3)
    off = 12
    xyz = np.zeros((nn, 3))
    num = np.zeros(nn, int)
    for i in range(nn):
        num[i] = np.frombuffer(data, np.int32, 1, off)[0]
        xyz[i] = np.frombuffer(data, np.float64, 3, off + 4)
        off += 28
    base = 1 if num[0] == 1 else 0
    rows = []
    for e in range(ne):
        _, nnpe = np.frombuffer(data, np.int32, 2, off)
        rows.append((np.frombuffer(data, np.int32, nnpe, off + 8) - base).tolist())
        off += 8 + 4 * nnpe
    return xyz[:, :abs(nd)], np.array(rows, np.int32), base
