"""Build recipe for libbeta_hip.so (gfx950) and the C++ host adapters.

Usage: python -m betanoox_amd.build   (or betanoox_amd.build.build_all())
Objects and the library go to betanoox_amd/_build/.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libbeta_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions stay disabled in every HIP compile.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]

SOURCES = ["capi.hip", "pattern.hip", "assemble.hip", "solver.hip", "plastic.hip", "multi.hip", "server.hip", "heat.hip", "formats.hip", "pack.hip", "forces.hip", "project.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def _compile(src):
    obj = os.path.join(OUT, os.path.splitext(os.path.basename(src))[0] + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_hip(jobs=8):
    os.makedirs(OUT, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_host():
    """C++ adapters (LargeMatrix/BasicAssembler-shaped) and their demo driver."""
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cpp")] if os.path.isdir(HOST) else []
    if not srcs:
        return None
    exe = os.path.join(OUT, "beta_hip_driver")
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + HOST] + srcs +
         ["-L" + OUT, "-lbeta_hip", "-Wl,-rpath,$ORIGIN", "-o", exe])
    return exe


def build_all():
    lib = build_hip()
    exe = build_host()
    return lib, exe


if __name__ == "__main__":
    print(build_all())
    sys.exit(0)
