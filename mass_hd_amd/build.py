"""Build recipe of the HIP library (libmass_lyo.so) and the C++ host program.

Everything is compiled in-tree with hipcc for gfx950:
  - lyo_fp64.hip   FP64 sweeps, -ffp-contract=off (IEEE-exact reference order)
  - lyo_fp32.hip   FP32 sweeps, FMA contraction allowed
  - lyo_capi.cpp   the C ABI of include/mass_lyo.h
  - lyo_config.cpp host-side initial configuration (std::mt19937)
The host program (mass_hip) is linked against the library with g++.

A second copy of the library and of the host program, with AddressSanitizer
in the host code of the ABI units (lib/asan/, bin/asan/, plus the ABI cycle
program of tests/native/), lets the tests check the host side's memory: no
access outside what it owns and nothing left allocated once every handle is
destroyed.  Device code is not instrumented; it is the same as in the
regular build.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mass_hd_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIBDIR, "libmass_lyo.so")
HOSTBIN = os.path.join(BINDIR, "mass_hip")
ASAN_LIBDIR = os.path.join(LIBDIR, "asan")
ASAN_BINDIR = os.path.join(BINDIR, "asan")
ASAN_LIB = os.path.join(ASAN_LIBDIR, "libmass_lyo.so")
ASAN_HOSTBIN = os.path.join(ASAN_BINDIR, "mass_hip")
ASAN_CYCLE = os.path.join(ASAN_BINDIR, "abi_cycle")
NATIVE_TESTS = os.path.join(ROOT, "tests", "native")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANGXX = "/opt/rocm/llvm/bin/clang++"
ARCH = "gfx950"
COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-Wall", "-Wno-unused-function",
    # no packed FP32 instructions (v_pk_*_f32): some of their forms have given
    # wrong results on this GPU type, so every HIP unit is built without them
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]

UNITS = [
    ("lyo_fp64.hip", ["-ffp-contract=off"]),
    # FMA only within one source expression: the same arithmetic in every
    # kernel that inlines the step, so batched and single runs agree bitwise
    ("lyo_fp32.hip", ["-ffp-contract=on"]),
    # the one-kernel network step computes the features as lyo_fp32.hip does
    ("lyo_net_fused.hip", ["-ffp-contract=on"]),
    ("lyo_capi.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("lyo_batch.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("lyo_net.cpp", ["-ffp-contract=fast", "-x", "hip"]),
]
# Host units built with g++ and the reference's flags (-O3, no -march): the
# initial fields go through libm cos/sin exactly as the reference's build
# does (g++ -O3 merges cos/sin of one argument into one sincos call, which
# can differ from separate cos and sin calls in the last bit).
GXX_UNITS = ["lyo_config.cpp"]
GXX_FLAGS = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra"]
# AddressSanitizer for the host code only (each flag after -Xarch_host, so
# the device code is compiled exactly as in the regular build)
ASAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer",
             "-Xarch_host", "-g"]
# the units whose host code owns memory: handles, batches, the network
ASAN_UNITS = ("lyo_capi.cpp", "lyo_batch.cpp", "lyo_net.cpp")
JOBS = max(1, min(8, os.cpu_count() or 1))


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), out))
    if verbose and out.strip():
        print(out)


def _obj(src, sub=""):
    return os.path.join(OBJDIR, sub, os.path.splitext(src)[0] + ".o")


def _compile_all(cmds, verbose):
    """Runs independent compile commands, JOBS at a time."""
    with ThreadPoolExecutor(JOBS) as pool:
        for f in [pool.submit(_run, c, verbose) for c in cmds]:
            f.result()


def build_lib(verbose=False, asan=True):
    """libmass_lyo.so and, with asan, its host-sanitized copy in lib/asan."""
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(os.path.join(OBJDIR, "asan"), exist_ok=True)
    cmds, objs, asan_objs = [], [], []
    for src, extra in UNITS:
        cmds.append([HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", _obj(src)])
        objs.append(_obj(src))
        if asan and src in ASAN_UNITS:
            cmds.append([HIPCC] + COMMON + extra + ASAN_HOST +
                        ["-c", os.path.join(CSRC, src), "-o", _obj(src, "asan")])
            asan_objs.append(_obj(src, "asan"))
        else:
            asan_objs.append(_obj(src))
    for src in GXX_UNITS:
        cmds.append(["g++"] + GXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", _obj(src)])
        objs.append(_obj(src))
        asan_objs.append(_obj(src))
    _compile_all(cmds, verbose)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    if asan:
        # the sanitizer runtime is linked into the executables that load this copy
        os.makedirs(ASAN_LIBDIR, exist_ok=True)
        tmp = ASAN_LIB + ".tmp"
        _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + asan_objs,
             verbose)
        os.replace(tmp, ASAN_LIB)
    return LIB


def build_host(verbose=False):
    """C++ host program: Model registry, runcard parser, JSON writer, driver."""
    if not os.path.isdir(HOST):
        return None
    os.makedirs(BINDIR, exist_ok=True)
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    if not srcs:
        return None
    tmp = HOSTBIN + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-pthread", "-Wall", "-Wextra",
          "-I" + os.path.join(ROOT, "include"),
          "-I" + HOST, "-o", tmp] + srcs +
         ["-L" + LIBDIR, "-lmass_lyo", "-Wl,-rpath,$ORIGIN/../lib"], verbose)
    os.replace(tmp, HOSTBIN)
    return HOSTBIN


def build_checks(verbose=False):
    """bin/format_check: the JSON writer against the number stream it
    replaces (tests/native/format_check.cpp); bin/format_check_tsan: the same
    under ThreadSanitizer (host code, g++), for its parallel parts."""
    os.makedirs(BINDIR, exist_ok=True)
    src = os.path.join(NATIVE_TESTS, "format_check.cpp")
    jobs = []
    for name, flags in (("format_check", ["-O2"]),
                        ("format_check_tsan", ["-O1", "-g", "-fsanitize=thread"])):
        out = os.path.join(BINDIR, name)
        jobs.append((["g++", "-std=c++17", "-pthread", "-Wall", "-Wextra"] + flags +
                     ["-I" + HOST, "-o", out + ".tmp", src], out))
    _compile_all([c for c, _ in jobs], verbose)
    for _, out in jobs:
        os.replace(out + ".tmp", out)


def build_asan_programs(verbose=False):
    """bin/asan/mass_hip (the host program) and bin/asan/abi_cycle (every
    handle kind created, used and destroyed), both with AddressSanitizer and
    linked against lib/asan/libmass_lyo.so."""
    os.makedirs(ASAN_BINDIR, exist_ok=True)
    san = ["-fsanitize=address", "-fno-omit-frame-pointer", "-g", "-O1"]
    link = ["-L" + ASAN_LIBDIR, "-lmass_lyo", "-Wl,-rpath,$ORIGIN/../../lib/asan"]
    inc = ["-I" + os.path.join(ROOT, "include")]
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    jobs = [([CLANGXX, "-std=c++17", "-pthread", "-Wall"] + san + inc + ["-I" + HOST, "-o",
             ASAN_HOSTBIN + ".tmp"] + srcs + link, ASAN_HOSTBIN),
            # a C program, linked by the C++ driver so that operator new and
            # delete are the sanitizer's (their callers then appear in the
            # allocation stacks)
            ([CLANGXX, "-x", "c", "-std=c11", "-Wall", "-Wextra"] + san + inc +
             ["-o", ASAN_CYCLE + ".tmp", os.path.join(NATIVE_TESTS, "abi_cycle.c"), "-x", "none"] +
             link, ASAN_CYCLE)]
    _compile_all([c for c, _ in jobs], verbose)
    for _, out in jobs:
        os.replace(out + ".tmp", out)


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)


def build_all(verbose=False):
    build_lib(verbose)
    build_host(verbose)
    build_asan_programs(verbose)
    build_checks(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
