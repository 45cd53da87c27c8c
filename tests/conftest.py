import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Runcard of the 128^2 plain lyotropic case (FP64, npc=1): tau -> tauNem =
# tauIso, elastic K -> LL, phase-field A -> AA; CC, KK, zetaI, conc are set
# explicitly because the reference leaves them without defaults.
CONFIG0_PHYS = dict(GammaQ=0.1, GammaP=0.1, xi=0.7, tauNem=2.0, tauIso=2.0, rho=40.0,
                    friction=0.0, AA=0.01, CC=0.0, LL=0.04, KK=0.04, zeta=0.02, zetaI=0.0,
                    conc=0.5)
# director angle theta = 90deg + pi*(U - 1/2) ~ U[0, pi); Q = S(nn - I/2), S = 1
# (amplitude 1/2); phi = 0.5 + U(-0.05, 0.05)
CONFIG0_INIT = dict(config="random", seed=1, angle=90.0, noise=1.0, qamp=0.5, phinoise=0.05)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_present():
    try:
        from mass_hd_amd import lyo
        lyo.load_library()
    except Exception:
        return False
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return False
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) != 0:
        return False
    return n.value > 0


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_present():
        pytest.fail("GPU test run without a usable GPU or without the built library")
