"""CPU-only checks of the C ABI: the library loads (no GPU needed for dlopen)
and exports every entry point include/beta_hip.h declares; calls that need a
device fail with a status, not a crash, when no GPU is visible."""
import ctypes as C
import os
import re

import pytest

from betanoox_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdrs = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")]
    names = set()
    for h in hdrs:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(bhip_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(capi.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(capi.EXPORTS) == declared_symbols()


def test_version_string():
    lib = capi.load_library()
    assert b"gfx950" in lib.bhip_version()


def test_create_without_gpu_reports_status():
    """No GPU in this container: bhip_create must return a status code."""
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be visible here")
    lib = capi.load_library()
    h = C.c_void_p()
    st = lib.bhip_create(0, C.byref(h))
    assert st in (capi.ERR_HIP, capi.ERR_ARG)
    assert not h.value


def test_spmv_shape_table():
    """The tile shapes the tuner chooses from (no GPU needed)."""
    from betanoox_amd import capi
    shapes = capi.spmv_shapes()
    assert len(shapes) == 21
    assert [s["id"] for s in shapes] == list(range(21))
    assert shapes[0]["slices_per_block"] == 4 and shapes[0]["unroll"] == 4 and shapes[0]["nontemporal"] == 0
    triples = {(s["slices_per_block"], s["unroll"], s["nontemporal"]) for s in shapes[:18]}
    assert triples == {(a, b, c) for a in (4, 8, 16) for b in (2, 4, 8) for c in (0, 1)}
    # 18-20: runs of union slots (unroll 0), streaming loads
    assert [(s["slices_per_block"], s["unroll"], s["nontemporal"]) for s in shapes[18:]] == [(4, 0, 1), (8, 0, 1),
                                                                                             (16, 0, 1)]
