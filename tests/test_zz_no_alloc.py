"""The allocation-free entry point cpic_field_energies_noalloc against
cpic_energies.

The caller owns the device fields and the output (allocated here with
hipMalloc through ctypes, as a caller that cannot let the library allocate
would own them).  Two calls in a row, the second on a smaller grid, each
give the E and B energies of a context holding the same fields, bit for
bit (same kernels, same block count, same summation order), in FP32 and
FP64.
"""
import ctypes

import numpy as np
import pytest

from cabanapic_amd import Simulation, load_library

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """The few HIP runtime calls the test needs, through ctypes."""

    def __init__(self):
        load_library()  # the library's own HIP runtime
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def upload(self, dev, a):
        assert self.L.hipMemcpy(dev, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def download(self, a, dev):
        assert self.L.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), dev, a.nbytes, D2H) == 0

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self, p):
        assert self.L.hipFree(p) == 0


def _fields(nx, ny, nz, dtype, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (9, (nx + 2) * (ny + 2) * (nz + 2))).astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_noalloc_energies_match_context(dtype):
    hip = Hip()
    lib = load_library()
    grids = [(61, 47, 33), (9, 7, 5)]  # the second call is smaller
    fields = [_fields(*g, dtype, seed=3 + i) for i, g in enumerate(grids)]
    expect = []
    for g, f in zip(grids, fields):
        with Simulation(*g, 0, dtype) as sim:
            sim.upload_fields(f)
            expect.append(np.array(sim.energies()))
    dev_f = hip.malloc(fields[0].nbytes)  # room for the larger grid
    dev_out = hip.malloc(16)
    try:
        for g, f, e in zip(grids, fields, expect):
            hip.upload(dev_f, f)
            rc = lib.cpic_field_energies_noalloc(dev_f, np.dtype(dtype).itemsize, *g, dev_out,
                                                 None)
            assert rc == 0, lib.cpic_last_error()
            hip.sync()
            got = np.zeros(2)
            hip.download(got, dev_out)
            assert np.array_equal(got, e), (g, got, e)
        # invalid arguments are refused before any launch
        assert lib.cpic_field_energies_noalloc(None, 4, 4, 4, 4, dev_out, None) == 1
        assert b"null" in lib.cpic_last_error()
        assert lib.cpic_field_energies_noalloc(dev_f, 2, 4, 4, 4, dev_out, None) == 1
        assert lib.cpic_field_energies_noalloc(dev_f, 4, 0, 4, 4, dev_out, None) == 1
        assert b"extents" in lib.cpic_last_error()
    finally:
        hip.free(dev_f)
        hip.free(dev_out)
