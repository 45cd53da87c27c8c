"""The C ABI library loads without a GPU and exports every function that
include/*.h declares (no device calls are made here)."""
import ctypes
import glob
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "mass_hd_amd", "lib", "libmass_lyo.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mass_lyo_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_the_abi():
    names = declared_functions()
    for n in ("mass_lyo_create", "mass_lyo_step", "mass_lyo_get_field", "mass_lyo_sums",
              "mass_lyo_configure", "mass_lyo_destroy", "mass_lyo_batch_step"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    for n in sorted(declared_functions()):
        assert hasattr(lib, n), n
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB], text=True)
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert declared_functions() <= exported


def test_abi_version():
    lib = ctypes.CDLL(LIB)
    assert lib.mass_lyo_abi_version() == 2


def test_header_compiles_as_c():
    src = "#include \"mass_lyo.h\"\nint main(void){mass_lyo_params p; (void)p; return 0;}\n"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-x", "c", "-", "-fsyntax-only"], input=src, text=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    assert r.returncode == 0, r.stdout


def test_noalloc_sums_reject_bad_arguments_on_the_host():
    """The argument checks of mass_lyo_state_sums_noalloc run before any
    device call, so they are exercised here without a GPU."""
    lib = ctypes.CDLL(LIB)
    fn = lib.mass_lyo_state_sums_noalloc
    fn.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p,
                   ctypes.c_void_p]
    lib.mass_lyo_last_error.argtypes = [ctypes.c_void_p]
    lib.mass_lyo_last_error.restype = ctypes.c_char_p
    p = ctypes.c_void_p(4096)  # never dereferenced: every call below is refused first
    max_stride = (2 ** 31 - 1) // 9
    cases = [
        (7, p, 10, p, None, 10, "precision"),                    # unknown precision
        (0, None, 10, p, None, 10, "null"),                      # no populations
        (0, p, 10, p, p, 10, "compensation"),                    # FP64 with phic
        (1, p, 10, p, None, 0, "at least 1"),                    # empty state
        (1, p, 10, p, None, 11, "plane_stride"),                 # stride below n
        (1, p, max_stride + 1, p, None, 100, "too large"),       # beyond the scratch's limit
    ]
    for prec, f, stride, phi, phic, n, msg in cases:
        assert fn(prec, f, stride, phi, phic, n, 40.0, p, None) == 1
        assert msg in lib.mass_lyo_last_error(None).decode()
