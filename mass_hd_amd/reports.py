"""Precision and statistics reports of the GPU path.

precision_report: the same run in FP64 (reference arithmetic, bitwise the
CPU oracle's) and FP32 (shifted populations, compensated phi) on the GPU,
started from the same initial fields; for every field of the run (the state
f, Qxx, Qyx, phi and every derived frame field: n, u, the stresses, MU, H)
the max and rms error of FP32 against FP64 relative to the field's largest
FP64 magnitude, the drift of sum(phi) in both, and the diagnostics (u_rms,
defects, phi histogram) of both.

    python -m mass_hd_amd.reports precision --size 1024 --steps 10000 --npc 2
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import lyo

CONFIG0_PHYS = dict(GammaQ=0.1, GammaP=0.1, xi=0.7, tauNem=2.0, tauIso=2.0, rho=40.0,
                    friction=0.0, AA=0.01, CC=0.0, LL=0.04, KK=0.04, zeta=0.02, zetaI=0.0,
                    conc=0.5)
CONFIG0_INIT = dict(config="random", seed=1, angle=90.0, noise=1.0, qamp=0.5, phinoise=0.05)
# every field of a plain lyotropic run (state and derived frame fields)
FIELDS = [k for k, v in sorted(lyo.FIELDS.items(), key=lambda kv: kv[1]) if v < lyo.FIELDS["m"]]


def _rel_errors(a, b):
    d = np.abs(a - b).reshape(-1)
    scale = np.abs(b).max()
    scale = scale if scale > 0 else 1.0
    return float(d.max() / scale), float(math.sqrt(np.mean(d * d)) / scale)


def precision_report(size=1024, steps=10000, npc=2, chunks=10, phys=None, init=None,
                     progress=None):
    """Returns a dict with per-field errors at `chunks` evenly spaced times."""
    phys = dict(CONFIG0_PHYS if phys is None else phys)
    init = dict(CONFIG0_INIT if init is None else init)
    m64 = lyo.Lyotropic(size, size, npc=npc, precision=lyo.FP64, **phys)
    m32 = lyo.Lyotropic(size, size, npc=npc, precision=lyo.FP32, **phys)
    m64.configure(**init)
    m32.configure(**init)
    p0_64, p0_32 = m64.sums()[1], m32.sums()[1]
    out = {"size": size, "steps": steps, "npc": npc, "series": []}
    done = 0
    t0 = time.time()
    for c in range(chunks):
        n = (steps * (c + 1)) // chunks - done
        m64.step(n)
        m32.step(n)
        done += n
        rec = {"step": done}
        for f in FIELDS:
            mx, rms = _rel_errors(m32.get(f), m64.get(f))
            rec[f] = {"max_rel": mx, "rms_rel": rms}
        s64, s32 = m64.sums()[1], m32.sums()[1]
        rec["phi_drift_fp64"] = (s64 - p0_64) / p0_64
        rec["phi_drift_fp32"] = (s32 - p0_32) / p0_32
        d64, d32 = m64.diagnostics(), m32.diagnostics()
        rec["u_rms"] = [d64["u_rms"], d32["u_rms"]]
        rec["defects"] = [d64["defects_pos"] + d64["defects_neg"],
                          d32["defects_pos"] + d32["defects_neg"]]
        h64, h32 = (np.asarray(d["phi_hist"], dtype=float) for d in (d64, d32))
        rec["phi_hist_l1"] = float(np.abs(h64 / h64.sum() - h32 / h32.sum()).sum())
        out["series"].append(rec)
        if progress:
            progress("step %d: max_rel ff %.2e Q %.2e phi %.2e u %.2e  (%.0fs)" % (
                done, rec["ff"]["max_rel"], rec["QQxx"]["max_rel"], rec["phi"]["max_rel"],
                rec["ux"]["max_rel"], time.time() - t0))
    m64.close()
    m32.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["precision"])
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--npc", type=int, default=2)
    ap.add_argument("--chunks", type=int, default=10)
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    r = precision_report(a.size, a.steps, a.npc, a.chunks,
                         progress=lambda s: print(s, file=sys.stderr, flush=True))
    txt = json.dumps(r, indent=1)
    if a.out:
        open(a.out, "w").write(txt)
    print(txt)


if __name__ == "__main__":
    main()
