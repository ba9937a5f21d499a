"""firex_reduce_noalloc, the allocation-free variant of firex_reduce, called
through the C ABI and checked against firex_reduce on two calls in a row, the
second on a smaller grid.

The variant's block partials live in one array per GPU that every call
shares and that keeps the last call's contents; the second, smaller grid
launches fewer blocks (210 against 1024), so a read of stale partials from the
first call would change its sums.  The six sums use the same kernels in the
same order as firex_reduce and must be equal bit for bit; the two L2 norms
take their square root on the device instead of the host, both IEEE
correctly rounded, and are held to 2^-52 relative.
"""
import ctypes

import numpy as np
import pytest

from fire_amd import FireXHipSolver, REDUCTIONS, device_count, seeded_emitters
from fire_amd import solver as fs

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
L2 = {"residual_l2", "flow_divergence_l2"}


@pytest.fixture(scope="module")
def hip():
    if device_count() < 1:
        pytest.fail("no HIP device visible to the allocation-free reduction test")
    fs.load_library()
    # the HIP runtime libfirex_hip.so is linked against (already loaded)
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipSetDevice.argtypes = [ctypes.c_int]
    return lib


def reductions_both_ways(hip, dims, storage):
    print(f"no-alloc reduction: {dims} species={storage}", flush=True)
    g = FireXHipSolver(dims, tier="hd", pressure_iterations=24, correction_iterations=24, species_storage=storage)
    g.set_emitters(seeded_emitters(dims, 2, radius_range=(4.0, 8.0), seed=11), jitter=0.1, seed=11)
    g.step(DT, 6)
    want = {name: g.reduce(name) for name in REDUCTIONS}
    assert hip.hipSetDevice(0) == 0
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), 8 * len(REDUCTIONS)) == 0
    try:
        lib = fs.load_library()
        for name, which in REDUCTIONS.items():
            rc = lib.firex_reduce_noalloc(g._h, which, ctypes.c_void_p(dev.value + 8 * which))
            assert rc == 0, fs.load_library().firex_last_error()
        g.sync()
        got = np.zeros(len(REDUCTIONS), dtype=np.float64)
        assert hip.hipMemcpy(got.ctypes.data, dev, got.nbytes, 2) == 0  # device to host
    finally:
        hip.hipFree(dev)
    g.close()
    return want, {name: float(got[which]) for name, which in REDUCTIONS.items()}


def test_noalloc_reduction_equals_firex_reduce_twice(hip):
    for dims, storage in (((128, 192, 128), "f32"), ((40, 56, 24), "f16")):
        want, got = reductions_both_ways(hip, dims, storage)
        for name in REDUCTIONS:
            if name in L2:
                assert got[name] == pytest.approx(want[name], rel=2.0 ** -52, abs=0.0), (dims, name)
            else:
                assert got[name] == want[name], (dims, name, got[name], want[name])
        assert want["sum_fuel"] > 0 and want["residual_l2"] > 0


def test_noalloc_reduction_rejects_bad_arguments(hip):
    lib = fs.load_library()
    g = FireXHipSolver((40, 56, 24), tier="hd", pressure_iterations=8, correction_iterations=8)
    assert lib.firex_reduce_noalloc(g._h, 0, None) == fs.ERR_INVALID
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), 8) == 0
    try:
        assert lib.firex_reduce_noalloc(g._h, 99, dev) == fs.ERR_INVALID
    finally:
        hip.hipFree(dev)
    g.close()
