"""Each plausible transcription mistake in the oracle's step terms fails a pin.

For every mistake below, a copy of oracle/lyo_oracle.c with that one change
is compiled into a temporary directory and the oracle's pins
(test_oracle_beris_edwards.py, test_oracle_pins.py) are run against it in a
child process (LYO_ORACLE_LIB points the binding at the copy).  The check
passes when at least one pin fails.  The mistakes are single sign or factor
changes of the terms of reference src/models/lyotropic.cpp:196-221 and
:281-297, the ones that the single-mode and conservation pins alone cannot
see.  Nothing here touches the GPU.
"""
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "lyo_oracle.c")
HDR_DIR = os.path.join(ROOT, "oracle")
PINS = ["tests/test_oracle_beris_edwards.py", "tests/test_oracle_pins.py"]

MISTAKES = {
    "Q advection sign": ("- vx * dxQxx - vy * dyQxx", "+ vx * dxQxx - vy * dyQxx"),
    "Q advection sign (y)": ("- vx * dxQyx - vy * dyQyx", "- vx * dxQyx + vy * dyQyx"),
    "co-rotation sign": ("- 2 * vorticity * Qyx", "+ 2 * vorticity * Qyx"),
    "capillary normal stress sign": ("(1 - p) + .5 * P->KK * (dyPhi * dyPhi - dxPhi * dxPhi)",
                                     "(1 - p) - .5 * P->KK * (dyPhi * dyPhi - dxPhi * dxPhi)"),
    "capillary shear stress sign": ("(1 - p) - P->KK * dxPhi * dyPhi",
                                    "(1 - p) + P->KK * dxPhi * dyPhi"),
    "Frank normal stress sign": ("+ P->LL * (dyQxx * dyQxx", "- P->LL * (dyQxx * dyQxx"),
    "Frank shear stress factor": ("- 2 * P->LL * (dxQxx * dyQxx", "- 1 * P->LL * (dxQxx * dyQxx"),
    "bulk CC pressure factor": ("+ .5 * P->CC * term * term", "+ .25 * P->CC * term * term"),
    "alignment (Q + I) term": ("(Qxx + 1) * (2 * dxux - traceQL)",
                               "(Qxx - 1) * (2 * dxux - traceQL)"),
    "alignment shear term": ("+ 2 * Qyx * shear - expansion", "- 2 * Qyx * shear - expansion"),
    "alignment trace term": ("Qyx * (expansion - traceQL)", "Qyx * (expansion + traceQL)"),
    "alignment stress": ("((Qxx * Qxx - 1) * Hxx", "((Qxx * Qxx + 1) * Hxx"),
    "antisymmetric stress sign": ("sigmaA = 2 * (Qxx * Hyx - Qyx * Hxx)",
                                  "sigmaA = 2 * (Qyx * Hxx - Qxx * Hyx)"),
}


@pytest.fixture(scope="module")
def source():
    with open(SRC) as f:
        return f.read()


@pytest.mark.parametrize("name", sorted(MISTAKES))
def test_planted_mistake_fails_a_pin(name, source):
    old, new = MISTAKES[name]
    # the plain lyotropic node functions (go-or-grow's copies follow them)
    a = source.index("static void update_quantities_at_node(")
    b = source.index("static void growth_boundary_fields(")
    assert source[a:b].count(old) == 1, "the term to alter is not unique: %r" % old
    altered = source[:a] + source[a:b].replace(old, new) + source[b:]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "lyo_oracle.c")
        lib = os.path.join(d, "liblyo_oracle.so")
        with open(src, "w") as f:
            f.write(altered)
        # the flags of oracle/Makefile
        subprocess.check_call(["gcc", "-std=gnu99", "-O3", "-fopenmp", "-fPIC",
                               "-ffp-contract=off", "-I" + HDR_DIR, "-shared", "-o", lib, src,
                               "-lm"])
        env = dict(os.environ, LYO_ORACLE_LIB=lib)
        r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                            "-m", "not gpu"] + PINS, cwd=ROOT, env=env,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=600)
    tail = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else ""
    assert r.returncode == 1 and " failed" in tail and "error" not in tail, (name, tail)
