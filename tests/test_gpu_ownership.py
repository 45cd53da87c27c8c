"""Ownership at the C ABI: a handle owns every device buffer it allocates and
mass_lyo_destroy (mass_lyo_batch_destroy, mass_lyo_net_destroy) frees them
all.  Each cycle below creates, uses and destroys handles of every kind at
1024^2 (a leaked plane would be 4-8 MiB); once the runtime's own pools have
grown to their working size (code objects, queues, scratch), further cycles
must leave the device's free memory unchanged.
"""
import ctypes

import pytest

from mass_hd_amd import lyo
from tests.conftest import CONFIG0_INIT, CONFIG0_PHYS

pytestmark = pytest.mark.gpu

L = 1024


def free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def cycle():
    # FP32 run in surrogate mode (one-kernel network step, calibration, the
    # third state buffer set) with diagnostics, a frame and a checkpoint
    net = lyo.Net(seed=3)
    src = lyo.Lyotropic(64, 64, npc=2, precision="fp32", **CONFIG0_PHYS)
    src.configure(**CONFIG0_INIT)
    src.step(3)
    net.collect(src)
    net.train(3)
    src.close()
    m = lyo.Lyotropic(L, L, npc=2, precision="fp32", **CONFIG0_PHYS)
    m.configure(**CONFIG0_INIT)
    lyo.enable_surrogate(m, net, judge_every=2, tol=1.0)
    lyo.set_surrogate_calibration(m, 1)
    m.step(3)
    lyo.set_surrogate_kernels(m, 0)
    m.step(2)
    m.diagnostics()
    m.get("sigmaXY")
    img = m.checkpoint()
    m.restore(img)
    m.close()
    net.close()
    # FP64 run with the reference's loop kernels, a walled FP64 run
    d = lyo.Lyotropic(256, 256, npc=2, precision="fp64", **CONFIG0_PHYS)
    d.configure(**CONFIG0_INIT)
    d.set_loop_kernels(True)
    d.step(2)
    d.close()
    w = lyo.Lyotropic(L, 256, npc=1, precision="fp64", bc=4, **CONFIG0_PHYS)
    w.configure(**CONFIG0_INIT)
    w.step(2)
    w.close()
    # a batch of runs
    b = lyo.Batch([dict(CONFIG0_PHYS, LX=s, LY=s) for s in (128, 256, 512)], npc=2)
    for i in range(3):
        b.configure(i, **CONFIG0_INIT)
    b.step(4)
    b.close()


def test_destroy_frees_every_device_buffer(require_gpu):
    free = [free_bytes()]
    for _ in range(5):
        cycle()
        free.append(free_bytes())
    mib = [(free[0] - f) / 2**20 for f in free]
    print("device memory in use after each cycle (MiB):", ["%.1f" % x for x in mib])
    # the last three cycles: no growth at all (a leaked plane is 4-8 MiB)
    assert max(free[-3:]) - min(free[-3:]) < 2**20, mib
