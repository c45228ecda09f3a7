"""Writes matrix_debug_main.json: the known answers the reference's dense
matrix debug main prints beside its calls (beta/math/matrix.cpp, the
"A.Solve {{3,1},{.3333,1.6667}}" and "B.Invert = {{...}}" lines), with the
inputs it sets up above them.  Reads the reference source as text; run in a
checkout that has it:  python tests/golden/make_matrix_debug_main.py REF_ROOT
"""
import json
import os
import re
import sys


def _matrix(text):
    rows = re.findall(r"\{([^{}]*)\}", text)
    return [[float(v) for v in r.split(",")] for r in rows]


def _cramer(a, b):
    det = a[0][0] * a[1][1] - a[0][1] * a[1][0]
    return [(b[0] * a[1][1] - a[0][1] * b[1]) / det, (a[0][0] * b[1] - b[0] * a[1][0]) / det]


def main(ref_root):
    path = os.path.join(ref_root, "beta", "math", "matrix.cpp")
    lines = open(path).read().split("\n")
    found = {}
    for i, line in enumerate(lines, 1):
        m = re.search(r'"A\.Solve (\{\{.*\}\})"', line)
        if m:
            found["lu_factors"] = {"line": i, "expected": _matrix(m.group(1)[1:-1])}
        m = re.search(r'"B\.Invert = (\{\{.*\}\})"', line)
        if m:
            found["invert"] = {"line": i, "expected": _matrix(m.group(1)[1:-1])}
        m = re.search(r'"A = (\{\{.*\}\})"', line)
        if m and "A" not in found:
            found["A"] = {"line": i, "value": _matrix(m.group(1)[1:-1])}
        m = re.search(r'"B = (\{\{.*\}\})"', line)
        if m and "B" not in found:
            found["B"] = {"line": i, "value": _matrix(m.group(1)[1:-1])}
        m = re.search(r"double AA\[\d+\] = \{([^}]*)\}", line)
        if m and "rhs" not in found:
            found["rhs"] = {"line": i, "value": [float(v) for v in m.group(1).split(",")]}
    need = ("A", "B", "rhs", "lu_factors", "invert")
    missing = [k for k in need if k not in found]
    if missing:
        raise SystemExit("not found in %s: %s" % (path, missing))
    out = {
        "source": "beta/math/matrix.cpp (debug main)",
        "A": found["A"]["value"], "A_line": found["A"]["line"],
        "B": found["B"]["value"], "B_line": found["B"]["line"],
        # A.Solve(AA): LU factors of A in place; the first two entries of AA are the
        # right-hand side of the 2 x 2 system, whose solution is computed here by Cramer's rule
        "solve_rhs": found["rhs"]["value"][:2], "rhs_line": found["rhs"]["line"],
        "solve_x": _cramer(found["A"]["value"], found["rhs"]["value"][:2]),
        "lu_factors": found["lu_factors"]["expected"], "lu_line": found["lu_factors"]["line"],
        "invert": found["invert"]["expected"], "invert_line": found["invert"]["line"],
        "printed_digits": 5,
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "matrix_debug_main.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
