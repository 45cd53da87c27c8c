"""Host-side memory checks with AddressSanitizer and LeakSanitizer.

The build keeps a copy of the library whose ABI units (handles, batches, the
network) have sanitized host code (mass_hd_amd/lib/asan), the host program
linked against it (mass_hd_amd/bin/asan/mass_hip) and a C program that
creates, uses and destroys every kind of handle (tests/native/abi_cycle.c).
Each case runs one of them and requires a clean exit: any out-of-bounds or
use-after-free access, or any host allocation still live at exit whose
allocation stack does not pass through the ROCm runtime (its own allocations,
tests/native/lsan_rocm.supp), ends the process with the sanitizer's report
and exit status 23.  The ownership contract: the handle owns everything it
allocates, mass_lyo_destroy frees it (the device side is checked by
test_gpu_ownership.py).

On a GPU, ROCm's sanitizer runtime tracks the HSA runtime's device
allocations too, and after the leak check at exit, while the HIP and HSA
runtimes unload, it stops on an internal CHECK
(sanitizer_allocator_device.h: a quarantined device block recycled after the
HSA runtime is gone).  That teardown CHECK, and nothing else, is tolerated:
it comes after the leak report, and any error report counts.
"""
import os
import subprocess

import pytest

from tests.conftest import gpu_present
from tests.test_host_program import GROWTH_RUNCARD, RUNCARD

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mass_hd_amd", "bin", "asan")
CYCLE = os.path.join(BIN, "abi_cycle")
HOSTEXE = os.path.join(BIN, "mass_hip")
SUPP = os.path.join(ROOT, "tests", "native", "lsan_rocm.supp")
TEARDOWN = "CHECK failed: sanitizer_allocator_device.h"


def sanitized(args, timeout=300, cwd=None):
    if not os.path.exists(args[0]):
        pytest.fail("%s is not built (run __graft_entry__.build())" % args[0])
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:exitcode=23:protect_shadow_gap=0"
    env["LSAN_OPTIONS"] = "suppressions=%s:print_suppressions=0" % SUPP
    print("running", " ".join(args), flush=True)
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       env=env, timeout=timeout, cwd=cwd)
    err = r.stderr
    errors = [l for l in err.splitlines()
              if "ERROR: AddressSanitizer" in l or "ERROR: LeakSanitizer" in l
              or ("CHECK failed" in l and TEARDOWN not in l)]
    rc = r.returncode
    if rc == 23 and not errors and TEARDOWN in err:
        rc = 0   # the runtimes' teardown CHECK after a clean leak check
    return rc, r.stdout, err, bool(errors)


def test_abi_error_paths_without_device():
    """Without a GPU every create fails with MASS_LYO_EDEVICE, after the
    parameter checks; nothing leaks on those paths."""
    if gpu_present():
        pytest.skip("a GPU is present: the GPU cycle test covers these paths")
    rc, out, err, report = sanitized([CYCLE, "--no-device"])
    assert not report and rc == 0, err
    assert "error paths ok" in out


def test_leak_checker_is_live():
    """The same run with one buffer dropped on purpose: the leak checker
    reports it (4321 bytes) and the exit status says so, so a silent checker
    cannot pass the cases here."""
    if gpu_present():
        pytest.skip("a GPU is present: the no-device paths do not run")
    rc, out, err, report = sanitized([CYCLE, "--no-device", "--leak"])
    assert report and rc == 23, (rc, err)
    assert "4321 byte(s)" in err


@pytest.mark.parametrize("extra,code", [
    ("nsubsteps = 2\n", 1),        # rejected by the model
    ("precision = fp16\n", 1),     # bad value
    ("bogus = 1\n", 1),            # unknown key
])
def test_host_program_error_paths(tmp_path, extra, code):
    """The host program's runcard errors stop before the model reaches the
    GPU; they leak nothing."""
    card = tmp_path / "run.dat"
    card.write_text(RUNCARD + extra)
    rc, out, err, report = sanitized([HOSTEXE, str(card), "-o", str(tmp_path / "o"), "-f"])
    assert not report, err
    assert rc == code, (rc, err)


def test_host_program_help():
    rc, out, err, report = sanitized([HOSTEXE, "--help"])
    assert not report and rc == 0, err


@pytest.mark.gpu
def test_abi_cycles(require_gpu):
    """Two cycles of every handle kind (FP32 with timing, diagnostics and a
    checkpoint; FP64 with the loop kernels and a state import; a walled run;
    go-or-grow; lyotropic-with-division; a batch with the autotuner; the
    network, its training and surrogate mode in both kernel modes), then
    calls that must fail: every call returns what it should, and the
    sanitizers find nothing."""
    rc, out, err, report = sanitized([CYCLE, "2"])
    print(out)
    assert not report, err[-6000:]
    assert rc == 0, err[-6000:]
    assert out.count("network and surrogate ok") == 2 and "error paths ok" in out


CARDS = {
    "fp64": RUNCARD,
    "fp32": RUNCARD.replace("npc = 1", "npc = 2") + "precision = fp32\n",
    "walls": RUNCARD + "bc = 4\n",
    "go-or-grow": GROWTH_RUNCARD,
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CARDS))
def test_host_program_runs(tmp_path, require_gpu, name):
    """Runcards through the sanitized host program, frames written: a clean
    exit with the model destroyed at the end."""
    card = tmp_path / "run.dat"
    card.write_text(CARDS[name])
    outdir = tmp_path / "out"
    rc, out, err, report = sanitized([HOSTEXE, str(card), "-o", str(outdir), "-f"])
    assert not report, err[-4000:]
    assert rc == 0, err[-4000:]
    assert "frame20.json" in os.listdir(outdir)
