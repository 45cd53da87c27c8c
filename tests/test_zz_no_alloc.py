"""mass_lyo_state_sums_noalloc against mass_lyo_sums, through the C ABI.

The state of a run (taken from its checkpoint image: the raw device planes
f, Qxx, Qyx, phi and, in FP32, phi's compensation plane) is copied into
device memory that the test allocates; the entry point, which allocates
nothing, must give bitwise the sums mass_lyo_sums gives on the run itself.
Three calls in a row share the entry point's device scratch, each on a
smaller state than the one before (4.41 M, 78 k and 3 k nodes: 1,077, 20
and 1 partial sums), alternating precision.
"""
import ctypes

import numpy as np
import pytest

from mass_hd_amd import lyo
from tests.conftest import CONFIG0_INIT, CONFIG0_PHYS

pytestmark = pytest.mark.gpu

HEADER = 64  # bytes of the checkpoint header before the planes


def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def noalloc(lib):
    fn = lib.mass_lyo_state_sums_noalloc
    fn.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p,
                   ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return fn


def test_noalloc_sums_equal_the_runs_sums(require_gpu):
    lib = lyo.load_library()
    fn = noalloc(lib)
    H = hip()
    cases = [(2100, 2100, lyo.FP32), (300, 260, lyo.FP64), (64, 48, lyo.FP32)]
    for LX, LY, prec in cases:
        print("noalloc sums %dx%d precision %d" % (LX, LY, prec), flush=True)
        m = lyo.Lyotropic(LX, LY, npc=2, precision=prec, **CONFIG0_PHYS)
        m.configure(**CONFIG0_INIT)
        m.step(3)
        fs, ps = m.sums()
        img = m.checkpoint()
        m.close()
        N = LX * LY
        dt = np.float64 if prec == lyo.FP64 else np.float32
        nplanes = 12 if prec == lyo.FP64 else 13
        planes = np.frombuffer(img, dtype=dt, offset=HEADER, count=nplanes * N).copy()
        assert len(img) == HEADER + planes.nbytes
        d, out = ctypes.c_void_p(), ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(d), planes.nbytes) == 0
        assert H.hipMalloc(ctypes.byref(out), 16) == 0
        try:
            assert H.hipMemcpy(d, planes.ctypes.data, planes.nbytes, 1) == 0  # host to device
            sz = planes.itemsize
            phi = d.value + 11 * N * sz
            phic = d.value + 12 * N * sz if prec == lyo.FP32 else None
            rc = fn(prec, d, N, phi, phic, N, CONFIG0_PHYS["rho"], out, None)
            assert rc == 0, lib.mass_lyo_last_error(None)
            assert H.hipDeviceSynchronize() == 0
            res = np.zeros(2)
            assert H.hipMemcpy(res.ctypes.data, out, 16, 2) == 0  # device to host
        finally:
            H.hipFree(d)
            H.hipFree(out)
        assert res[0] == fs and res[1] == ps, (LX, LY, prec, res, fs, ps)
