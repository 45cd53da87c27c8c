"""Build recipe of the HIP library (libmass_lyo.so) and the C++ host program.

Everything is compiled in-tree with hipcc for gfx950:
  - lyo_fp64.hip   FP64 sweeps, -ffp-contract=off (IEEE-exact reference order)
  - lyo_fp32.hip   FP32 sweeps, FMA contraction allowed
  - lyo_capi.cpp   the C ABI of include/mass_lyo.h
  - lyo_config.cpp host-side initial configuration (std::mt19937)
The host program (mass_hip) is linked against the library with g++.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mass_hd_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIBDIR, "libmass_lyo.so")
HOSTBIN = os.path.join(BINDIR, "mass_hip")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
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


def build_lib(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        _run([HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    for src in GXX_UNITS:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        _run(["g++"] + GXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
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
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"),
          "-I" + HOST, "-o", tmp] + srcs +
         ["-L" + LIBDIR, "-lmass_lyo", "-Wl,-rpath,$ORIGIN/../lib"], verbose)
    os.replace(tmp, HOSTBIN)
    return HOSTBIN


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)


def build_all(verbose=False):
    build_lib(verbose)
    build_host(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
