"""The benchmark workloads of the lyotropic step on one MI355X.

  config0  128^2 FP64, npc=1, 2000 steps (seed 1; reference init contract)
  config1  1024^2, npc=2, 10^4 steps: FP64 vs FP32 error report
  config2  64 independent runs (L in {128,192,256,384,512}, zeta log-uniform in
           [1e-3, 5e-2], xi in {0.5, 0.7, 1.0}), FP32, npc=2, one batch
  config3  surrogate corrector at 512^2 (trained on config2's 512^2 members)
  config4  4096^2 FP32 plain npc=2 vs surrogate; defects and u_rms against
           the FP64 plain run

    python -m mass_hd_amd.workloads config2 --steps 5000 --out r.json
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import lyo
from .reports import CONFIG0_INIT, CONFIG0_PHYS, precision_report

SIZES = [128, 192, 256, 384, 512]
XIS = [0.5, 0.7, 1.0]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def config2_runs(n=64, sampler_seed=2024, base_seed=1000):
    """The sweep: sizes, activities and alignment parameters per run."""
    rng = np.random.default_rng(sampler_seed)
    runs = []
    for i in range(n):
        L = int(rng.choice(SIZES))
        zeta = float(math.exp(rng.uniform(math.log(1e-3), math.log(5e-2))))
        xi = float(rng.choice(XIS))
        runs.append(dict(CONFIG0_PHYS, LX=L, LY=L, zeta=zeta, xi=xi, seed=base_seed + i))
    return runs


def make_batch(runs, npc=2, precision="fp32", autotune=True):
    b = lyo.Batch([{k: v for k, v in r.items() if k != "seed"} for r in runs], npc=npc,
                  precision=precision)
    tuned = b.autotune() if autotune else None
    for i, r in enumerate(runs):
        b.configure(i, **dict(CONFIG0_INIT, seed=r["seed"]))
    return b, tuned


def config2(steps=5000, n=64, sampler_seed=2024, base_seed=1000, chunks=5):
    runs = config2_runs(n, sampler_seed, base_seed)
    nodes = sum(r["LX"] * r["LY"] for r in runs)
    b, tuned = make_batch(runs)
    log("config2: %d runs, %d nodes, autotuned launch %s" % (n, nodes, tuned))
    t_ms = 0.0
    done = 0
    for c in range(chunks):
        k = steps * (c + 1) // chunks - done
        t_ms += b.time_steps(k)
        done += k
        log("  %d steps, %.1f MLUPS" % (done, nodes * done / (t_ms / 1e3) / 1e6))
    diag = [b.diagnostics(i) for i in range(n)]
    out = {"runs": len(runs), "nodes": nodes, "steps": steps, "ms": t_ms,
           "mlups": nodes * steps / (t_ms / 1e3) / 1e6, "autotuned": tuned,
           "sampler_seed": sampler_seed, "base_seed": base_seed,
           "u_rms": [d["u_rms"] for d in diag],
           "defects": [d["defects_pos"] + d["defects_neg"] for d in diag],
           "fsum_ok": all(abs(b.sums(i)[0] - 40.0 * r["LX"] * r["LY"]) < 1.0
                          for i, r in enumerate(runs))}
    b.close()
    return out


def train_surrogate(frames_per_run=8, spacing=50, warm=200, iters=2000, lr=2e-3, n=64,
                    sampler_seed=2024, base_seed=1000, seed=0, matched=None):
    """Trains the surrogate on frames of the 512^2 members of the config2
    sweep (their own plain npc=2 steps), or -- matched = (zeta, xi, count) --
    of `count` 512^2 runs with the physics of the run it will stand in for
    (other seeds than that run's)."""
    if matched:
        zeta, xi, count = matched
        runs = [dict(CONFIG0_PHYS, LX=512, LY=512, zeta=zeta, xi=xi, seed=base_seed + i)
                for i in range(count)]
    else:
        runs = [r for r in config2_runs(n, sampler_seed, base_seed) if r["LX"] == 512]
    net = lyo.Net(seed=seed)
    b, _ = make_batch(runs, autotune=False)
    b.step(warm)
    for _ in range(frames_per_run):
        b.step(spacing)
        for i in range(len(runs)):
            net.collect(lyo.RunRef(b, i))
    pairs = net.dataset_size * 512 * 512
    log("surrogate: %d frames (%d predictor->corrector pairs) from %d runs" %
        (net.dataset_size, pairs, len(runs)))
    t0 = time.time()
    l0, l1 = net.train(iters, lr)
    log("  trained %d iterations in %.1fs: loss %.4g -> %.4g" % (iters, time.time() - t0, l0, l1))
    b.close()
    return net, {"frames": net.dataset_size, "pairs": pairs, "runs": len(runs), "iters": iters,
                 "warm": warm, "spacing": spacing, "loss_first": l0, "loss_last": l1}


def compare(L, net, steps, zeta=None, seed=7, judge_every=50, tol=1e-3, every=None,
            fp64_steps=0, twin=True, calibrate=False, twin_eps=1e-9):
    """Plain npc=2 vs surrogate (and FP64 plain over fp64_steps) from one
    initial state; time series of u_rms and defect density.  twin: a second
    plain run from the same state with phi perturbed by twin_eps (random
    signs), whose distance from the plain run is the spread of two plain runs
    once their trajectories have parted -- the yardstick for the surrogate,
    whose trajectory parts from the plain one within its first steps (at
    4096^2 a 1e-9 perturbation takes ~6e4 steps to show, 1e-3 far fewer)."""
    phys = dict(CONFIG0_PHYS)
    if zeta is not None:
        phys["zeta"] = zeta
    every = every or max(1, steps // 20)
    series = {}
    timing = {}
    modes = (["fp64"] if fp64_steps else []) + ["plain"] + (["twin"] if twin else []) + ["surrogate"]
    for mode in modes:
        prec = "fp64" if mode == "fp64" else "fp32"
        m = lyo.Lyotropic(L, L, npc=2, precision=prec, **phys)
        m.configure(**dict(CONFIG0_INIT, seed=seed))
        if mode == "twin":
            sgn = np.sign(np.random.default_rng(seed).uniform(-1, 1, L * L))
            m.set_state(m.get("ff"), m.get("QQxx"), m.get("QQyx"),
                        m.get("phi") + twin_eps * sgn.reshape(m.get("phi").shape))
        if mode == "surrogate":
            lyo.enable_surrogate(m, net, judge_every, tol)
            lyo.set_surrogate_calibration(m, 1 if calibrate else 0)
        n = fp64_steps if mode == "fp64" else steps
        rows, t_ms, done = [], 0.0, 0
        while done < n:
            k = min(every, n - done)
            t0 = time.time()
            m.step(k)
            m.synchronize()
            t_ms += (time.time() - t0) * 1e3
            done += k
            d = m.diagnostics()
            rows.append((done, d["u_rms"], d["defect_density"]))
            log("  %s %d: u_rms %.4g, defect density %.4g (%.0f MLUPS)" %
                (mode, done, d["u_rms"], d["defect_density"], L * L * done / (t_ms / 1e3) / 1e6))
        timing[mode] = {"steps": n, "ms": t_ms, "mlups": L * L * n / (t_ms / 1e3) / 1e6}
        if mode == "surrogate":
            timing[mode]["stats"] = lyo.surrogate_stats(m)
            timing[mode]["calibration"] = lyo.surrogate_calibration(m)
        series[mode] = rows
        log("%s %d^2: %d steps, %.0f MLUPS, final u_rms %.4g, defect density %.4g" %
            (mode, L, n, timing[mode]["mlups"], rows[-1][1], rows[-1][2]))
        m.close()
    return {"L": L, "steps": steps, "series": series, "timing": timing,
            "relative_to_plain": relative_stats(series, windows=10)}


def relative_stats(series, windows=10):
    """Relative differences of u_rms and defect density from the plain run:
    per checkpoint (max and mean of |d|) and over `windows` equal windows of
    checkpoints (the statistics of a chaotic flow are its time averages)."""
    base = np.array(series["plain"])
    out = {}
    for mode, rows in series.items():
        if mode == "plain":
            continue
        r = np.array(rows)
        n = min(len(r), len(base))
        if n == 0:
            continue
        r, b = r[:n], base[:n]
        rec = {"checkpoints": n}
        for col, name in ((1, "u_rms"), (2, "defect_density")):
            d = r[:, col] / b[:, col] - 1.0
            w = max(1, n // windows) if mode != "fp64" else n
            nw = n // w
            wm = [float(r[k * w:(k + 1) * w, col].mean() / b[k * w:(k + 1) * w, col].mean() - 1.0)
                  for k in range(nw)]
            rec[name] = {"checkpoint_max_abs": float(np.abs(d).max()),
                         "checkpoint_mean": float(d.mean()),
                         "checkpoint_rms": float(np.sqrt(np.mean(d * d))),
                         "window_steps": int(w * (r[1, 0] - r[0, 0]) if n > 1 else r[0, 0]),
                         "window_means": wm,
                         "window_max_abs": float(np.abs(wm).max()) if wm else None}
        out[mode] = rec
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["config1", "config2", "config3", "config4"])
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--iters", type=int, default=0,
                    help="training iterations (0: 2000 for config3, 8000 for config4)")
    ap.add_argument("--fp64-steps", type=int, default=10000)
    ap.add_argument("--every", type=int, default=1000)
    ap.add_argument("--out", default="")
    # training frames: config3 (2,000 steps) from the first 600 steps of the
    # 512^2 runs; config4 (10^5 steps) from steps 1,500-5,000, the developed
    # flow (frames from the first 600 steps only let the long 4096^2 run drift:
    # u_rms +26%, defect density +33% at 5x10^4 steps, against +0.4% / +4%)
    ap.add_argument("--warm", type=int, default=0,
                    help="plain steps of the training runs before their first frame "
                         "(0: 200 for config3, 1000 for config4)")
    ap.add_argument("--spacing", type=int, default=0,
                    help="steps between training frames (0: 50 for config3, 500 for config4)")
    ap.add_argument("--frames", type=int, default=0,
                    help="training frames per run (0: 8 for config3, 24 for config4; more "
                         "frames and iterations lowered config4's long-run bias)")
    ap.add_argument("--matched", type=int, default=0,
                    help="N > 0: train on N 512^2 runs with the run's own zeta and xi instead "
                         "of the config2 sweep's 512^2 members")
    ap.add_argument("--calibrate", type=int, default=0,
                    help="1: the surrogate's judged steps recalibrate its output scales")
    ap.add_argument("--twin-eps", type=float, default=1e-9,
                    help="amplitude of the twin run's phi perturbation (1e-9 stays on the "
                         "plain trajectory for most of config4; 1e-3 decorrelates early)")
    ap.add_argument("--no-twin", action="store_true",
                    help="skip the plain run from a state perturbed by 1e-9")
    a = ap.parse_args(argv)
    if a.what == "config1":
        r = precision_report(1024, a.steps or 10000, 2, 10, progress=log)
    elif a.what == "config2":
        r = config2(a.steps or 5000)
    else:
        late = a.what == "config4"
        matched = None
        if a.matched > 0:
            matched = (0.03 if late else CONFIG0_PHYS["zeta"], CONFIG0_PHYS["xi"], a.matched)
        net, tr = train_surrogate(frames_per_run=a.frames or (24 if late else 8),
                                  spacing=a.spacing or (500 if late else 50),
                                  warm=a.warm or (1000 if late else 200),
                                  iters=a.iters or (8000 if late else 2000),
                                  matched=matched)
        tr["matched"] = matched
        if a.what == "config3":
            r = {"training": tr, "run": compare(512, net, a.steps or 2000, twin=not a.no_twin,
                                                twin_eps=a.twin_eps,
                                                calibrate=bool(a.calibrate))}
        else:
            r = {"training": tr, "run": compare(4096, net, a.steps or 100000, zeta=0.03, seed=7,
                                                every=a.every, fp64_steps=a.fp64_steps,
                                                twin=not a.no_twin, calibrate=bool(a.calibrate),
                                                twin_eps=a.twin_eps)}
    txt = json.dumps(r, indent=1)
    if a.out:
        open(a.out, "w").write(txt)
    print(txt)


if __name__ == "__main__":
    main()
