"""K / b exchange files (no GPU): the library's writers against the
reference's own SparseMatrixUtility.cpp (compiled unchanged into oracle/_ref)
and against the stream formats the reference's RCV and vector dumps use.

  - CSR ASCII and CSR binary files are byte-identical to the reference's
    writeCSRMatrixInASCII / writeCSRMatrixInBinary on the same arrays
  - the reference's readCSRMatrixInASCII / readCSRMatrixInBinary and
    readLoadVector read the library's files back exactly
  - RCV lower/upper text: type line, "L nnz", then per row "\\t i j value"
    with the value in setw(24) setprecision(16) of a default-format stream
    (printf "%24.16g"), one line per row
  - LoadVector.txt / Solution.txt: scientific setw(24) setprecision(16)
    (printf "%24.16e"), the solution file led by its count
"""
import numpy as np
import pytest

from betanoox_amd import capi
from oracle import oracle as O


def _random_upper(n, seed):
    """A symmetric pattern stored as its upper triangle, 1-based (PARDISO
    storage), diagonal first in every row, values spanning many magnitudes."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        cols = sorted({i} | set(rng.choice(np.arange(i, n), size=min(4, n - i), replace=False).tolist()))
        rows.append(cols)
    ia = np.zeros(n + 1, np.int32)
    for i, c in enumerate(rows):
        ia[i + 1] = ia[i] + len(c)
    ja = np.concatenate([np.array(c, np.int32) for c in rows])
    a = rng.standard_normal(ja.shape[0]) * 10.0 ** rng.integers(-12, 12, ja.shape[0])
    a[ia[:-1]] = np.abs(a[ia[:-1]]) + 1.0
    a[3] = 0.0
    a[5] = -1.0 / 3.0
    return ia + 1, ja + 1, a


def _need_ref():
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (make -C oracle ref needs the reference sources)")


@pytest.mark.parametrize("binary", [False, True])
def test_csr_files_byte_identical_to_reference(tmp_path, binary):
    _need_ref()
    ia, ja, a = _random_upper(37, 5)
    fmt = capi.FMT_CSR_BINARY if binary else capi.FMT_CSR_ASCII
    mine, ref = tmp_path / "mine", tmp_path / "ref"
    capi.write_csr_file(mine, fmt, ia, ja, a)
    O.ref_write_csr(ref, binary, ia, ja, a)
    assert mine.read_bytes() == ref.read_bytes()
    ria, rja, ra = O.ref_read_csr(mine, binary)
    assert np.array_equal(ria, ia) and np.array_equal(rja, ja)
    assert np.array_equal(ra, a)


@pytest.mark.parametrize("fmt,typ", [(capi.FMT_RCV_LOWER, "1"), (capi.FMT_RCV_UPPER, "2")])
def test_rcv_text_layout(tmp_path, fmt, typ):
    ia, ja, a = _random_upper(9, 2)
    ia0, ja0 = ia - 1, ja - 1
    path = tmp_path / "k.txt"
    capi.write_csr_file(path, fmt, ia0, ja0, a)
    text = path.read_text()
    expect = typ + "\n" + "%d %d\n" % (9, ja.shape[0])
    for i in range(9):
        expect += "".join("\t%d %d %24.16g" % (i, ja0[k], a[k]) for k in range(ia0[i], ia0[i + 1])) + "\n"
    assert text == expect


@pytest.mark.parametrize("kind", [capi.VEC_LOAD, capi.VEC_SOLUTION])
def test_vector_files(tmp_path, kind):
    v = np.array([1.0, -2.5e-300, 3.0e300, 0.0, 1.0 / 3.0, -7.0])
    path = tmp_path / "v.txt"
    capi.write_vector(path, kind, v)
    expect = ("%d\n" % v.shape[0] if kind == capi.VEC_SOLUTION else "") + "".join("%24.16e\n" % x for x in v)
    assert path.read_text() == expect
    assert np.array_equal(capi.read_vector(path, kind, v.shape[0]), v)
    if kind == capi.VEC_LOAD and O.ref_lib() is not None:
        assert np.array_equal(O.ref_read_load_vector(path, v.shape[0]), v)


def test_write_csr_file_rejects_bad_arguments(tmp_path):
    lib = capi.load_library()
    assert lib.bhip_write_csr_file(str(tmp_path / "x").encode(), 99, 0, 0,
                                   np.zeros(1, np.int32).ctypes.data_as(capi._P_I), None, None) == capi.ERR_ARG
