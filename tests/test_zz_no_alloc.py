"""hamsx_factor_noalloc (the factorisation that allocates nothing, its
workspace in a fixed per-GPU array) against hamsx_factor on the same
matrices, called through the C-ABI twice in a row -- N = 6,320 (outer-block
path, last panel and last outer block partial), then N = 432 -- with
bit-for-bit equal factors, pivots and INFO; and its size limit."""
import ctypes as C

import numpy as np
import pytest

from hamsx import _lib
from hamsx.mesh import combine, cylinder
from hamsx.solver import FrequencySolver

pytestmark = pytest.mark.gpu


def pair(na, nz, nr):
    bodies = [cylinder(4.0, 8.0, na, nz, nr, -15.0), cylinder(3.0, 6.0, na, nz, nr, 15.0)]
    return combine(bodies, [np.array([-15.0, 0, 0]), np.array([15.0, 0, 0])])


class DeviceBuffer:
    """Device memory for the test's own inputs and outputs, from the HIP
    runtime the library itself is linked against (the test may allocate;
    the entry point under test does not)."""
    hip = None

    def __init__(self, nbytes):
        if DeviceBuffer.hip is None:
            DeviceBuffer.hip = C.CDLL("libamdhip64.so.7")
        self.nbytes = nbytes
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0

    def upload(self, a):
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def download(self, a):
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(a.nbytes), 2) == 0

    def free(self):
        self.hip.hipFree(self.ptr)


def test_noalloc_factor_equals_context_factor():
    lib = _lib.load()
    for m, depth in ((pair(79, 35, 5), -1.0), (pair(24, 6, 3), 40.0)):
        n = m.nelem
        print(f"[gpu] hamsx_factor_noalloc n={n} depth={depth}", flush=True)
        with FrequencySolver(m) as s:
            s.calwaveproperts(1.2, depth)
            s.assemble()
            A = s.get_matrix("A")
            info = s.factor()
            LU, ipiv = s.get_matrix("LU")
        a_host = np.asfortranarray(A)
        bufs = [DeviceBuffer(a_host.nbytes), DeviceBuffer(4 * n), DeviceBuffer(4)]
        try:
            a, piv, inf = bufs
            a.upload(a_host)
            inf.upload(np.array([-7], dtype=np.int32))
            rc = lib.hamsx_factor_noalloc(n, a.ptr, piv.ptr, inf.ptr, None)
            assert rc == 0, lib.hamsx_last_error()
            assert DeviceBuffer.hip.hipDeviceSynchronize() == 0
            LU2 = np.zeros_like(a_host)
            piv2 = np.zeros(n, dtype=np.int32)
            info2 = np.zeros(1, dtype=np.int32)
            a.download(LU2)
            piv.download(piv2)
            inf.download(info2)
        finally:
            for b in bufs:
                b.free()
        np.testing.assert_array_equal(LU2, LU)
        np.testing.assert_array_equal(piv2, ipiv)
        assert int(info2[0]) == info == 0


def test_noalloc_factor_rejects_sizes_beyond_its_workspace():
    lib = _lib.load()
    big = _lib.HAMSX_NOALLOC_MAX_N + 1
    assert lib.hamsx_factor_noalloc(big, C.c_void_p(8), C.c_void_p(8), C.c_void_p(8), None) == _lib.HAMSX_EARG
    assert b"HAMSX_NOALLOC_MAX_N" in lib.hamsx_last_error()
