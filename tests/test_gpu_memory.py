"""Device memory of the C ABI's temporaries is returned.

Calls that allocate scratch inside the library (the copy and read stream
benchmarks, an assembly with a state vector, a heat run, the lower-triangle
export) are repeated after a warm-up call of each; the device's free memory
afterwards is where it was (within 64 MiB, the allocator's rounding).
Leaking the temporaries would lose over 2 GB here.
"""
import pytest

from betanoox_amd import capi, meshgen, setup_problem

pytestmark = pytest.mark.gpu


def _free():
    """Free device bytes from the HIP runtime the library itself loaded."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    assert hip.hipDeviceSynchronize() == 0
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_temporaries_released(hip_ctx_factory):
    p = meshgen.diffusion_cube(40)
    ctx = setup_problem(hip_ctx_factory(), p)
    u0 = p.extra["u0"]
    o = capi.Context.opts(tol=1e-8, precision=capi.PREC_FP32_IR, inner_tol=1e-3, maxit=20000)

    def calls():
        ctx.assemble(capi.ASM_HEAT_TRANSIENT, coef=[1.0, p.extra["dt"]], u_state=u0)
        ctx.heat_run(2, u0, capi.HEAT_X0_PREVIOUS, opts=o)
        ctx.values_lower()
        ctx.bench_stream(1 << 28, 2)
        ctx.bench_read(1 << 28, 2)

    calls()  # the context's own buffers and graphs are allocated here
    f0 = _free()
    for _ in range(3):
        calls()
    f1 = _free()
    assert f0 - f1 < (64 << 20), (f0, f1)
