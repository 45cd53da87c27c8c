"""Benchmark: MLUPS of the lyotropic active-nematic step at 4096^2 on one MI355X.

Workload (the large single-domain case): 4096 x 4096 periodic D2Q9, FP32 with
shifted populations and compensated phi, npc = 2 (predictor + 2 correctors),
zeta = 0.03, init seed 7 (random director, phi = 0.5 + U(-0.05, 0.05)),
200 warm-up + 500 timed steps.  The same grid is also timed in FP64 (the
reference arithmetic) with the fused kernels and with one kernel per
reference loop, in FP64 as a no-slip box (bc = 4, ghost layer), and in
surrogate mode (predictor + one corrector + the network, judged every 50
steps); go-or-grow is timed on the reference's own 500 x 500 proliferation
runcard.

MLUPS = LX*LY*steps / seconds / 1e6; one step = predictor + npc correctors
(with npc <= 2 the whole step is one kernel).  The timed region is bracketed
by HIP events on the library's stream; the fields are resident in HBM
throughout.

roofline: the kernel with the longest mean launch.  Its algorithmic bytes are
the per-sweep figure of the step's data flow -- one sweep reads and writes the
model state and the predictor's half-step state once, 36 values per node
(predictor: read f, Q, phi = 12, write f*, fn*, Q+D, Q+D/2, phi terms = 24;
corrector: read 24, write 12) -- times the sweeps one launch performs: 1 for
a predictor or corrector sweep, 2 for the fused last corrector pair, 1 + npc
for the one-kernel step.  In FP32 that is 144 B per node per sweep against
8 TB/s.  The fused kernels keep intermediate results in registers and LDS, so
they move fewer distinct bytes than their sweeps' worth:
"distinct_bytes_per_launch" counts every plane a launch touches once (one-
kernel step: read and write f', Q, phi, phic = 26 values; pair: 38;
predictor: 24), "traffic" is the HBM bytes the counters measured per launch
and "valu_busy_frac" the counted VALU instructions' issue time over the
launch time (profiles/).

cpu_baseline: the CPU oracle (the reference's algorithm and layout, OpenMP)
on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PHYS = dict(GammaQ=0.1, GammaP=0.1, xi=0.7, tauNem=2.0, tauIso=2.0, rho=40.0, friction=0.0,
            AA=0.01, CC=0.0, LL=0.04, KK=0.04, zeta=0.03, zetaI=0.0, conc=0.5)
INIT = dict(config="random", seed=7, angle=90.0, noise=1.0, qamp=0.5, phinoise=0.05)
HBM_PEAK = 8.0e12
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_fp32_4096_npc2.json")
# algorithmic values per node per launch (x4 bytes in FP32)
SWEEPS = {"predictor": 1, "corrector": 1, "pair": 2, "step": 3}   # sweeps per launch
SWEEP_VALUES = 36                                        # values per node per sweep
DISTINCT = {"predictor": 24, "corrector": 37, "pair": 38, "step": 26}  # FP32, npc=2 data flow
KERNEL = {"predictor": "sweep_kernel<float, false, 0, false>",
          "corrector": "sweep_kernel<float, false, 1, true>",
          "pair": "pair_kernel<float, false, true>",
          "step": "full_kernel<float>"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def gpu_run(lyo, L, precision, npc, steps, warmup, chunk=0, pair_chunk=0, fuse=-1, alt=-1,
            full=-1, step_chunk=0, wpb=0, loops=False, bc=0):
    m = lyo.Lyotropic(L, L, npc=npc, precision=precision, bc=bc, **PHYS)
    m.set_launch(chunk, wpb)
    m.set_schedule(fuse, pair_chunk, alt, full, step_chunk)
    if loops:
        m.set_loop_kernels(True)
    m.configure(**INIT)
    if warmup:
        m.step(warmup)
    m.synchronize()
    total_ms, sweeps = m.time_steps(steps, per_sweep=True)
    m.runtime_checks()
    m.close()
    return total_ms, dict(zip(("predictor", "corrector", "pair", "step"), sweeps))


# the reference's go-or-grow proliferation runcard
# (cases/20may/proliferation_p2_bulk/alpha0p00100_pc1p20/run.dat)
GROWTH_PHYS = dict(GammaQ=0.3, GammaP=5e-2, xi=0.3, tauNem=1.0, tauIso=1.0, rho=40.0,
                   friction=0.0, AA=0.03, CC=0.1, LL=0.002, KK=0.06, zeta=0.0, zetaI=0.0,
                   conc=1.0)
GROWTH_PARAMS = dict(B=0.1, Snem=1.0, relax_steps=2000, relax_dt=1.0, relax_phi=1, relax_Q=1,
                     chi_config="noise", chi0=1.0, chi_noise=0.0, chi_length=0.0, Dchi=0.0,
                     Achi=0.0, Ochi=0.0, phiswitch=0.0, growTogether=0, beta=0.0,
                     death_rate=0.0, death_time=1.0, death_radius=0.0, division_rate=0.1,
                     division_time=2.0, division_radius=3.0, alpha=0.001, phi_critical=1.2)
GROWTH_INIT = dict(config="circle", seed=1001, radius=60.0, level=250.0, noise=1.0, angle=0.0)


def growth_leg(lyo, steps, L=500):
    """go-or-grow at 500^2, npc=1 (the reference's proliferation runcard):
    configure (2000 dry relaxation steps, reset, phenotype), `steps` untimed,
    then `steps` timed steps, whose masks are redrawn every third step from
    phi read back to the host."""
    m = lyo.Lyotropic(L, L, npc=1, precision=lyo.FP64, **GROWTH_PHYS)
    m.set_growth(**GROWTH_PARAMS)
    t0 = time.time()
    m.configure(**GROWTH_INIT)
    m.synchronize()
    t_conf = time.time() - t0
    m.step(steps)
    m.synchronize()
    total_ms, _ = m.time_steps(steps, per_sweep=False)
    phi = m.get("phi")
    m.close()
    return {"workload": "go-or-grow 500x500 periodic, npc=1, FP64 (proliferation runcard)",
            "mlups": round(L * L * steps / (total_ms / 1e3) / 1e6, 1),
            "ms_per_step": round(total_ms / steps, 4), "configure_s": round(t_conf, 3),
            "phi_sum_after": round(float(phi.sum()), 3)}


def surrogate_leg(lyo, L, steps, judge_every=50, tol=1e-3, train_iters=500, warmup=200):
    """Surrogate corrector at LxL against the plain npc=2 step: a network
    trained briefly on frames of a 512^2 plain run (the full training is
    `python -m mass_hd_amd.workloads config3`), then `steps` surrogate steps
    timed with HIP events (one in judge_every of them a judged plain step),
    against `steps` plain steps of the same run after `warmup` steps."""
    net = lyo.Net(seed=0)
    src = lyo.Lyotropic(512, 512, npc=2, precision=lyo.FP32, **PHYS)
    src.configure(**INIT)
    src.step(200)
    for _ in range(8):
        src.step(25)
        net.collect(src)
    src.close()
    l0, l1 = net.train(train_iters, lr=2e-3)
    m = lyo.Lyotropic(L, L, npc=2, precision=lyo.FP32, **PHYS)
    m.configure(**INIT)
    m.step(warmup)
    m.synchronize()
    t_plain, _ = m.time_steps(steps, per_sweep=False)
    lyo.enable_surrogate(m, net, judge_every, tol)
    # a judged step and a surrogate step; the timed steps then hold one
    # judged step per judge_every steps
    m.step(2)
    m.synchronize()
    t_sur, _ = m.time_steps(steps, per_sweep=False)
    st = lyo.surrogate_stats(m)
    m.close()
    net.close()
    N = L * L
    return {"surrogate_mlups": round(N * steps / (t_sur / 1e3) / 1e6, 1),
            "plain_mlups_same_run": round(N * steps / (t_plain / 1e3) / 1e6, 1),
            "surrogate_over_plain": round(t_plain / t_sur, 3),
            "judge_every": judge_every, "tol": tol,
            "judged_velocity_error_last": st["last_error"],
            "judged_velocity_error_max": st["max_error"],
            "surrogate_steps": st["surrogate_steps"], "fallbacks": st["fallbacks"],
            "net_flop_per_lu": 2 * 9 * (8 * 32 + 32 * 32 + 32 * 32 + 32 * 8),
            "train_loss": [round(l0, 4), round(l1, 5)]}


def cpu_baseline(L, steps, threads, bc=0):
    from oracle import pyoracle as po
    o = po.Oracle(L, L, npc=2, nthreads=threads, bc=bc, **PHYS)
    o.configure(**INIT)
    o.step(1)  # warm-up (first touch of the work arrays)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    return L * L * steps / dt / 1e6


def cpu_growth(steps, threads, relax_steps=200, L=500):
    """The oracle on the go-or-grow runcard (sample: 200 instead of 2,000
    relaxation steps in configure, then `steps` timed steps; the mask draws
    are serial as in the reference)."""
    from oracle import pyoracle as po
    o = po.Oracle(L, L, npc=1, nthreads=threads, **GROWTH_PHYS)
    o.set_growth(**dict(GROWTH_PARAMS, relax_steps=relax_steps))
    o.configure(**GROWTH_INIT)
    o.step(1)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    return L * L * steps / dt / 1e6


def calibrate(gib=1, reps=10):
    """Plain device-to-device copies of a known size through the HIP runtime:
    the practical HBM ceiling, and the reference workload for scaling the
    FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = gib << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
        hip.hipMemset(p, 0, ctypes.c_size_t(n))
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # warm-up
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(b, a, ctypes.c_size_t(n), 3, None)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    gbps = 2.0 * n * reps / (ms.value / 1e3) / 1e9
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibration": "device-to-device copy", "bytes_read_per_copy": n,
                      "bytes_written_per_copy": n, "copies": reps + 1,
                      "copy_GBps_read_plus_write": round(gbps, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--calibrate", action="store_true",
                    help="time plain 1 GiB device copies (HBM ceiling, PMC scale)")
    ap.add_argument("--gpus", type=int, default=1)
    # 200 + 500 steps: the throughput of a long run (the first few dozen steps
    # after start-up run ~8% slower at 4096^2)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--npc", type=int, default=2)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--pair-chunk", type=int, default=0)
    ap.add_argument("--no-fuse", action="store_true", help="separate corrector sweeps")
    ap.add_argument("--no-full", action="store_true",
                    help="predictor + fused pair instead of the one-kernel step")
    ap.add_argument("--step-chunk", type=int, default=0)
    ap.add_argument("--wpb", type=int, default=0, help="waves per block (0: default)")
    ap.add_argument("--alt", type=int, default=-1,
                    help="march order: 0 every chunk upwards, 1 alternating (default)")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-fp64-loops", action="store_true",
                    help="skip the FP64 run with one kernel per reference loop")
    ap.add_argument("--skip-walls", action="store_true",
                    help="skip the FP64 no-slip box (bc=4) leg")
    ap.add_argument("--skip-growth", action="store_true",
                    help="skip the go-or-grow leg (500^2 proliferation runcard)")
    ap.add_argument("--growth-steps", type=int, default=300)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--skip-surrogate", action="store_true")
    ap.add_argument("--surrogate-steps", type=int, default=50,
                    help="timed steps of the surrogate leg (50: one judged step among them)")
    ap.add_argument("--surrogate-train", type=int, default=500,
                    help="Adam iterations of the surrogate leg's training")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.calibrate:
        calibrate()
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path "
              "(the x-slab decomposition is design only)", file=sys.stderr)
        sys.exit(2)

    from mass_hd_amd import lyo
    L, npc = args.size, args.npc
    N = L * L

    fuse = 0 if args.no_fuse else -1
    alt = args.alt
    extra = {}
    # go-or-grow first: after a sustained 4096^2 leg the next few hundred ms of
    # this launch-bound run are 25-30% slower
    if not args.skip_growth:
        log("go-or-grow 500x500: configure + %d + %d timed steps" % (args.growth_steps,
                                                                    args.growth_steps))
        extra["go_or_grow"] = growth_leg(lyo, args.growth_steps)
        log("go-or-grow: %s" % extra["go_or_grow"])

    log("fp32 %dx%d npc=%d: %d warmup + %d timed steps" % (L, L, npc, args.warmup, args.steps))
    t32, k32 = gpu_run(lyo, L, lyo.FP32, npc, args.steps, args.warmup, args.chunk,
                       args.pair_chunk, fuse, alt, 0 if args.no_full else -1, args.step_chunk,
                       args.wpb)
    mlups32 = N * args.steps / (t32 / 1e3) / 1e6
    log("fp32: %.3f ms/step, %.0f MLUPS, mean launch ms: %s"
        % (t32 / args.steps, mlups32, {k: round(v, 4) for k, v in k32.items() if v}))

    if not args.skip_fp64:
        s64 = max(2, args.steps // 4)
        log("fp64 %dx%d npc=%d: %d timed steps" % (L, L, npc, s64))
        w64 = max(2, args.warmup // 4)
        t64, k64 = gpu_run(lyo, L, lyo.FP64, npc, s64, w64, args.chunk, 0, fuse, alt,
                           0 if args.no_full else -1, args.step_chunk, args.wpb)
        mlups64 = N * s64 / (t64 / 1e3) / 1e6
        log("fp64: %.3f ms/step, %.0f MLUPS" % (t64 / s64, mlups64))
        extra.update({"fp64_mlups": round(mlups64, 1),
                      "fp32_over_fp64": round(mlups32 / mlups64, 3),
                      "fp64_kernel_set": "fused strip-march kernels (one-kernel step for "
                                         "npc=2), reference expression order",
                      "fp64_launch_ms": {k: round(v, 4) for k, v in k64.items() if v}})
        if not args.skip_fp64_loops:
            log("fp64 loop kernels %dx%d npc=%d: %d timed steps" % (L, L, npc, s64))
            tl, _ = gpu_run(lyo, L, lyo.FP64, npc, s64, w64, loops=True)
            mlupsl = N * s64 / (tl / 1e3) / 1e6
            log("fp64 loop kernels: %.3f ms/step, %.0f MLUPS" % (tl / s64, mlupsl))
            extra.update({"fp64_loops_mlups": round(mlupsl, 1),
                          "fp32_over_fp64_loops": round(mlups32 / mlupsl, 3),
                          "fp64_loops_kernel_set": "one kernel per reference node loop "
                                                   "(UpdateQuantities storing 23 derived "
                                                   "fields, UpdateFields, Move), SoA planes"})
        if not args.skip_walls:
            log("fp64 no-slip box (bc=4) %dx%d npc=%d: %d timed steps" % (L, L, npc, s64))
            tw, _ = gpu_run(lyo, L, lyo.FP64, npc, s64, w64, bc=4)
            mlupsw = N * s64 / (tw / 1e3) / 1e6
            log("fp64 walls: %.3f ms/step, %.0f MLUPS" % (tw / s64, mlupsw))
            extra.update({"fp64_walls_bc4_mlups": round(mlupsw, 1),
                          "fp64_walls_kernel_set": "ghost layer, three fused tile sweeps "
                                                   "(conditions on Q, phi, u, MU, sigma "
                                                   "in LDS; LB kernels; Move in the "
                                                   "correctors' reads)"})
            if not args.skip_fp64_loops:
                twl, _ = gpu_run(lyo, L, lyo.FP64, npc, s64, w64, loops=True, bc=4)
                mlupswl = N * s64 / (twl / 1e3) / 1e6
                log("fp64 walls, loop kernels: %.3f ms/step, %.0f MLUPS" % (twl / s64, mlupswl))
                extra["fp64_walls_bc4_loops_mlups"] = round(mlupswl, 1)

    dom = max((k for k in k32 if k32[k] > 0), key=lambda k: k32[k])
    # the one-kernel step performs all 1 + npc sweeps of a step
    sweeps = (1 + npc) if dom == "step" else SWEEPS[dom]
    alg_bytes = sweeps * SWEEP_VALUES * 4 * N
    achieved = alg_bytes / (k32[dom] / 1e3)
    distinct = DISTINCT[dom] * 4 * N
    traffic = valu_ms = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("size") == L and t.get("npc") == npc and dom in t.get("kernels", {}):
            traffic = t["kernels"][dom].get("bytes_per_launch")
            valu_ms = t["kernels"][dom].get("valu_issue_ms")

    if not args.skip_surrogate:
        log("surrogate %dx%d: training on 512^2 frames, %d timed steps" % (L, L, args.surrogate_steps))
        extra["surrogate"] = surrogate_leg(lyo, L, args.surrogate_steps,
                                           train_iters=args.surrogate_train,
                                           warmup=args.warmup)
        log("surrogate: %s" % extra["surrogate"])

    cpu = None
    if not args.skip_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log("cpu baseline: oracle %dx%d npc=2, %d threads, 1 warmup + %d steps"
            % (L, L, threads, args.cpu_steps))
        v = cpu_baseline(L, args.cpu_steps, threads)
        cpu = {"value": round(v, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
               "sample": "%dx%d FP64 npc=2, %d steps after 1 warm-up step, OpenMP" %
                         (L, L, args.cpu_steps)}
        log("cpu: %.2f MLUPS" % v)
        # the reference's CPU path beside the walls and go-or-grow legs
        if "fp64_walls_bc4_mlups" in extra:
            vw = cpu_baseline(L, args.cpu_steps, threads, bc=4)
            extra["cpu_walls_bc4"] = {"value": round(vw, 2), "unit": "MLUPS", "cores": threads,
                                      "kind": "port", "sample": "%dx%d FP64 bc=4 npc=2, %d steps "
                                      "after 1 warm-up step, OpenMP" % (L, L, args.cpu_steps)}
            log("cpu walls: %.2f MLUPS" % vw)
        if "go_or_grow" in extra:
            vg = cpu_growth(20, threads)
            extra["cpu_go_or_grow"] = {"value": round(vg, 2), "unit": "MLUPS", "cores": threads,
                                       "kind": "port", "sample": "500x500 go-or-grow runcard with "
                                       "200 relaxation steps, 20 steps after 1 warm-up step, "
                                       "OpenMP loops, serial mask draws"}
            log("cpu go-or-grow: %.2f MLUPS" % vg)

    line = {
        "metric": "MLUPS at %d^2 (FP32 plain, npc=%d)" % (L, npc),
        "value": round(mlups32, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (random director field, seed 7)",
        "config": {"workload": "lyotropic active nematic %dx%d periodic D2Q9, FP32 shifted "
                               "populations, npc=%d, zeta=0.03" % (L, L, npc),
                   "LX": L, "LY": L, "npc": npc},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": traffic,
                     "kernel": KERNEL[dom], "kernel_ms": round(k32[dom], 4),
                     "alg_bytes_per_launch": alg_bytes,
                     "distinct_bytes_per_launch": distinct,
                     "step_alg_GBps": round((1 + npc) * SWEEP_VALUES * 4 * N * args.steps
                                            / (t32 / 1e3) / 1e9, 1),
                     "launch_ms": {k: round(v, 4) for k, v in k32.items() if v},
                     "traffic_GBps": (round(traffic / (k32[dom] / 1e3) / 1e9, 1)
                                      if traffic else None),
                     # the fused step is bound by VALU issue: profiled VALU
                     # instructions x 4 cycles over the measured launch time
                     "valu_busy_frac": (round(valu_ms / k32[dom], 3) if valu_ms else None)},
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
