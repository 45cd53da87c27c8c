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

roofline: the kernel with the longest mean launch (the one-kernel step by
default).  Its algorithmic bytes are the compulsory HBM traffic of one launch:
the model state read once and written once, 12 values per node each way
(f' 9, Qxx, Qyx, phi), 96 B per node in FP32.  The compensation term of phi
(4 B each way) is overhead, not algorithm.  "per_sweep_equivalent" keeps the
per-sweep figure of the step's data flow (36 values per node per sweep,
3 sweeps at npc = 2: 432 B per node), which counts the predictor's
half-step state that the fused kernel keeps in registers and LDS.
"traffic" is the HBM bytes per launch that the rocprofv3 counters measured
(profiles/traffic_fp32_4096_npc2.json).
"device_copy_GBps" is the rate of plain 1 GiB device-to-device copies measured
in the same run, the practical streaming ceiling for read + write traffic.
"valu" gives the kernel's counted VALU instructions (SQ_INSTS_VALU) at 4
cycles per wave64 instruction on 256 CUs x 4 SIMDs at the 2.4 GHz peak engine
clock, against the measured launch time: the one-kernel step runs close to
both its VALU issue time and the copy rate.

Leg order: the FP32 headline run is configured first; the FP64 leg then runs
on its own handle (its warm-up long enough for the engine clock to settle
after the load step: a rocprofv3 GRBM_GUI_ACTIVE trace shows the clock
dipping to ~1.6 GHz for ~30 ms after work starts on an idle GPU), and the
headline's warm-up and timed steps follow with no idle gap in between.

cpu_baseline: the CPU oracle (the reference's algorithm and AoS layout,
built with the reference's -O3 -fopenmp) in a child process on the host: one
thread (the reference's default) at 1024^2 and the GPU's share of the host
cores (16, or fewer if fewer are available) at the headline size, each
5 warm-up + 20 timed steps.
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
MFMA_BF16_PEAK = 2.5e15                          # dense BF16 FLOP/s
VALU_PEAK_INSTR_PER_S = 256 * 4 * 2.4e9 / 4     # wave64 VALU instructions per second
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_fp32_4096_npc2.json")
LAUNCHES = ("predictor", "corrector", "pair", "step")
# compulsory values per node a launch must read and write (FP32 x4 bytes):
# the one-kernel step reads and writes the state (f' 9, Q 2, phi); the
# predictor reads the state and writes its outputs (f* 9, Q* 2, dphi); a
# corrector or the fused pair reads both and writes the next state
COMPULSORY = {"predictor": 24, "corrector": 36, "pair": 36, "step": 24}
SWEEPS = {"predictor": 1, "corrector": 1, "pair": 2, "step": 3}   # sweeps per launch
SWEEP_VALUES = 36                                        # values per node per sweep
KERNEL = {"predictor": "sweep_kernel<float, false, 0, false>",
          "corrector": "sweep_kernel<float, false, 1, true>",
          "pair": "pair_kernel<float, false, true>",
          "step": "full_kernel<float, false, 2>"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_run(lyo, L, precision, npc, chunk=0, pair_chunk=0, fuse=-1, alt=-1, full=-1,
             step_chunk=0, wpb=0, loops=False, bc=0):
    m = lyo.Lyotropic(L, L, npc=npc, precision=precision, bc=bc, **PHYS)
    m.set_launch(chunk, wpb)
    m.set_schedule(fuse, pair_chunk, alt, full, step_chunk)
    if loops:
        m.set_loop_kernels(True)
    m.configure(**INIT)
    m.synchronize()
    return m


def timed(m, steps, warmup):
    """warmup untimed steps, then `steps` steps between HIP events on the
    run's stream; returns (total ms, mean launch ms per launch kind)"""
    if warmup:
        m.step(warmup)
    m.synchronize()
    total_ms, sweeps = m.time_steps(steps, per_sweep=True)
    m.runtime_checks()
    return total_ms, dict(zip(LAUNCHES, sweeps))


def gpu_run(lyo, L, precision, npc, steps, warmup, **kw):
    m = make_run(lyo, L, precision, npc, **kw)
    try:
        return timed(m, steps, warmup)
    finally:
        m.close()


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


# the same proliferation runcard as the host program reads it
RUNCARD = """model = go-or-grow
config = circle
nsteps = {steps}
ninfo = 100
LX = 500
LY = 500
bc = 0
seed = 1001
GammaQ = 0.3
GammaP = 5e-2
zeta = 0
zetaI = 0
friction = 0.0
tauNem = 1.0
tauIso = 1.0
AA = 0.03
CC = 0.1
LL = 0.002
KK = 0.06
rho = 40
xi = 0.3
B = 0.1
Snem = 1
level = 250
radius = 60
conc = 1
noise = 1
angle = 0
relax-steps = 2000
relax-dt = 1.0
relax-phi = 1
relax-Q = 1
chi-config = noise
chi0 = 1
chi-noise = 0
chi-length = 0
Dchi = 0
Achi = 0
Ochi = 0
phiswitch = 0
growTogether = 0
beta = 0
death-rate = 0
death-time = 1
death-radius = 0
division-rate = 0.1
division-time = 2
division-radius = 3
alpha = 0.001
phi-critical = 1.2
"""


def runcard_leg(steps):
    """The proliferation runcard end to end through the host program
    (mass_hd_amd/bin/mass_hip, a child process): configure, `steps` steps and
    a JSON frame every 100 steps (each ~6.5 M numbers, 70 MB), as a MASS user
    runs it; the run's own report of its total and output-writing time."""
    import re
    import shutil
    import subprocess
    import tempfile
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mass_hd_amd", "bin", "mass_hip")
    tmp = tempfile.mkdtemp(prefix="mass_runcard_")
    try:
        card = os.path.join(tmp, "run.dat")
        with open(card, "w") as f:
            f.write(RUNCARD.format(steps=steps))
        t0 = time.time()
        r = subprocess.run([exe, card, "-o", os.path.join(tmp, "out") + "/", "-f"],
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600)
        wall = time.time() - t0
        if r.returncode != 0:
            raise RuntimeError("host program failed: %s" % r.stderr[-2000:])
        frames = len([f for f in os.listdir(os.path.join(tmp, "out")) if f.startswith("frame")])
        total = float(re.search(r"Total run time :\s*([0-9.]+)", r.stdout).group(1))
        write = float(re.search(r"Total time spent writing output :\s*([0-9.]+)", r.stdout).group(1))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return {"workload": "go-or-grow 500x500 proliferation runcard through mass_hip, npc=1, FP64, "
                        "a JSON frame every 100 steps",
            "steps": steps, "frames": frames, "run_s": total, "write_s": write,
            "process_wall_s": round(wall, 2)}


def checkpoint_leg(m, reps=3):
    """Binary checkpoint of the run after its timed steps, through the C ABI:
    the image (64-byte header + the raw device planes) saved into a host
    buffer (pageable, touched once before) and loaded back, best of `reps`;
    the run continues from the loaded image bitwise (tests)."""
    import ctypes
    import numpy as np
    n = ctypes.c_uint64()
    m._check(m.lib.mass_lyo_checkpoint_size(m.h, ctypes.byref(n)))
    size = int(n.value)
    buf = np.ones(size, dtype=np.uint8)
    ptr = ctypes.c_void_p(buf.ctypes.data)
    save, load = [], []
    for _ in range(reps):
        m.synchronize()
        t0 = time.perf_counter()
        m._check(m.lib.mass_lyo_checkpoint_save(m.h, ptr, n.value))
        t1 = time.perf_counter()
        m._check(m.lib.mass_lyo_checkpoint_load(m.h, ptr, n.value))
        m.synchronize()
        t2 = time.perf_counter()
        save.append(t1 - t0)
        load.append(t2 - t1)
    return {"workload": "checkpoint of the %dx%d FP32 run" % (m.LX, m.LY), "bytes": size,
            "save_ms": round(1e3 * min(save), 2), "load_ms": round(1e3 * min(load), 2),
            "save_GBps": round(size / min(save) / 1e9, 2),
            "load_GBps": round(size / min(load) / 1e9, 2)}


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
    # the same run with the network part as separate feature, convolution and
    # increment kernels (features and activations through HBM)
    lyo.set_surrogate_kernels(m, 0)
    m.step(2)
    m.synchronize()
    t_sep, _ = m.time_steps(steps, per_sweep=False)
    m.close()
    net.close()
    # the structural bound of a surrogate step: the predictor + one corrector
    # it still runs (the npc = 1 one-kernel step, timed on its own) plus the
    # network's FLOPs at the dense BF16 MFMA peak
    m1 = lyo.Lyotropic(L, L, npc=1, precision=lyo.FP32, **PHYS)
    m1.configure(**INIT)
    m1.step(warmup)
    m1.synchronize()
    t_one, _ = m1.time_steps(steps, per_sweep=False)
    m1.close()
    N = L * L
    flop = 2 * 9 * (8 * 32 + 32 * 32 + 32 * 32 + 32 * 8)
    net_ms = flop * N / MFMA_BF16_PEAK * 1e3
    bound_ms = t_one / steps + net_ms
    return {"surrogate_mlups": round(N * steps / (t_sur / 1e3) / 1e6, 1),
            "plain_mlups_same_run": round(N * steps / (t_plain / 1e3) / 1e6, 1),
            "surrogate_over_plain": round(t_plain / t_sur, 3),
            "judge_every": judge_every, "tol": tol,
            "judged_velocity_error_last": st["last_error"],
            "judged_velocity_error_max": st["max_error"],
            "surrogate_steps": st["surrogate_steps"], "fallbacks": st["fallbacks"],
            "net_flop_per_lu": flop,
            "ms_per_step": round(t_sur / steps, 4),
            "plain_ms_per_step": round(t_plain / steps, 4),
            "network_kernels": "one kernel: features, the four layers (activations in "
                               "registers), the increment",
            "separate_kernels_ms_per_step": round(t_sep / steps, 4),
            # the network's FLOPs over the time the surrogate step spends
            # beyond the predictor + one corrector (the one-kernel network
            # step and the amortised judge): a lower bound on its MFMA rate;
            # counters in profiles/surrogate_fp32_4096.md
            "network_tflops_beyond_step": round(
                flop * N / max(1e-9, (t_sur - t_one) / steps / 1e3) / 1e12, 1),
            "network_frac_of_bf16_peak_beyond_step": round(
                flop * N / max(1e-9, (t_sur - t_one) / steps / 1e3) / MFMA_BF16_PEAK, 4),
            "structural_bound": {
                "npc1_step_ms": round(t_one / steps, 4),
                "network_ms_at_bf16_peak": round(net_ms, 4),
                "bound_ms": round(bound_ms, 4),
                "best_possible_over_plain": round(t_plain / steps / bound_ms, 3),
                "note": "predictor + one corrector (still run) + the network at 2.5 PFLOP/s "
                        "dense BF16: the surrogate cannot beat the fused plain step on this GPU"},
            "train_loss": [round(l0, 4), round(l1, 5)]}


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_available": avail}


def cpu_child(spec):
    """Runs in a child process (so the OpenMP placement variables take
    effect): times the oracle on one workload and prints its MLUPS."""
    from oracle import pyoracle as po
    L, threads = spec["L"], spec["threads"]
    if spec["kind"] == "growth":
        o = po.Oracle(L, L, npc=1, nthreads=threads, **GROWTH_PHYS)
        o.set_growth(**dict(GROWTH_PARAMS, relax_steps=spec["relax_steps"]))
        o.configure(**GROWTH_INIT)
    else:
        o = po.Oracle(L, L, npc=spec["npc"], nthreads=threads, bc=spec["bc"], **PHYS)
        o.configure(**INIT)
    o.step(spec["warmup"])
    t0 = time.perf_counter()
    o.step(spec["steps"])
    dt = time.perf_counter() - t0
    o.close()
    print(json.dumps({"mlups": L * L * spec["steps"] / dt / 1e6, "seconds": dt}), flush=True)


def cpu_time(kind, L, steps, warmup, threads, npc=2, bc=0, relax_steps=0):
    """The oracle in a child process: threads = 0 is the reference's serial
    default (OpenMP off), otherwise OpenMP with OMP_PLACES=cores and
    OMP_PROC_BIND=close."""
    import subprocess
    spec = dict(kind=kind, L=L, steps=steps, warmup=warmup, threads=threads, npc=npc, bc=bc,
                relax_steps=relax_steps)
    env = dict(os.environ, OMP_PLACES="cores", OMP_PROC_BIND="close",
               OMP_NUM_THREADS=str(max(1, threads)))
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", json.dumps(spec)],
                       env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=900)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline child failed: %s" % r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1])["mlups"]


def calibrate(gib=1, reps=10, quiet=False):
    """Plain device-to-device copies of a known size through the HIP runtime:
    the practical HBM ceiling, and the reference workload for scaling the
    FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc.  Returns the copy
    rate in GB/s (bytes read + written)."""
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
    if not quiet:
        print(json.dumps({"calibration": "device-to-device copy", "bytes_read_per_copy": n,
                          "bytes_written_per_copy": n, "copies": reps + 1,
                          "copy_GBps_read_plus_write": round(gbps, 1)}), flush=True)
    return gbps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--calibrate", action="store_true",
                    help="time plain 1 GiB device copies (HBM ceiling, PMC scale)")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--gpus", type=int, default=1)
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
    ap.add_argument("--skip-host", action="store_true",
                    help="skip the proliferation runcard through the host program")
    ap.add_argument("--runcard-steps", type=int, default=2000)
    ap.add_argument("--skip-checkpoint", action="store_true",
                    help="skip the checkpoint save/load timing of the FP32 run")
    ap.add_argument("--fp64-preload-ms", type=float, default=80.0,
                    help="minimum untimed FP64 work before its timed steps (clock settling)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-warmup", type=int, default=5)
    ap.add_argument("--cpu-size-1t", type=int, default=1024,
                    help="grid of the single-thread CPU sample")
    ap.add_argument("--skip-surrogate", action="store_true")
    ap.add_argument("--surrogate-steps", type=int, default=50,
                    help="timed steps of the surrogate leg (50: one judged step among them)")
    ap.add_argument("--surrogate-train", type=int, default=500,
                    help="Adam iterations of the surrogate leg's training")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the multi-core CPU sample (0: min(16, available))")
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(json.loads(args.cpu_child))
        return
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
    sched = dict(chunk=args.chunk, pair_chunk=args.pair_chunk, fuse=0 if args.no_fuse else -1,
                 alt=args.alt, full=0 if args.no_full else -1, step_chunk=args.step_chunk,
                 wpb=args.wpb)
    extra = {}
    # go-or-grow first: it is launch-bound, and right after a sustained 4096^2
    # leg its first few hundred ms run 25-30% slower
    if not args.skip_growth:
        log("go-or-grow 500x500: configure + %d + %d timed steps" % (args.growth_steps,
                                                                    args.growth_steps))
        extra["go_or_grow"] = growth_leg(lyo, args.growth_steps)
        log("go-or-grow: %s" % extra["go_or_grow"])
    if not args.skip_host:
        log("proliferation runcard through the host program: %d steps" % args.runcard_steps)
        extra["host_runcard"] = runcard_leg(args.runcard_steps)
        log("runcard: %s" % extra["host_runcard"])

    log("fp32 %dx%d npc=%d: configure" % (L, L, npc))
    m32 = make_run(lyo, L, lyo.FP32, npc, **sched)
    m64 = None
    if not args.skip_fp64:
        s64 = max(2, args.steps // 4)
        m64 = make_run(lyo, L, lyo.FP64, npc, **sched)
        # untimed FP64 steps until the engine clock has settled, then the
        # timed FP64 steps; the FP32 steps follow right after
        t_one, _ = m64.time_steps(1, per_sweep=False)
        w64 = max(2, args.warmup // 4, int(args.fp64_preload_ms / max(t_one, 1e-3)) + 1)
        log("fp64 %dx%d npc=%d: %d warm-up + %d timed steps" % (L, L, npc, w64, s64))
        t64, k64 = timed(m64, s64, w64)
        mlups64 = N * s64 / (t64 / 1e3) / 1e6
        log("fp64: %.3f ms/step, %.0f MLUPS" % (t64 / s64, mlups64))
    log("fp32 %dx%d npc=%d: %d warmup + %d timed steps" % (L, L, npc, args.warmup, args.steps))
    t32, k32 = timed(m32, args.steps, args.warmup)
    if not args.skip_checkpoint:
        extra["checkpoint"] = checkpoint_leg(m32)
        log("checkpoint: %s" % extra["checkpoint"])
    m32.close()
    if m64 is not None:
        m64.close()
    mlups32 = N * args.steps / (t32 / 1e3) / 1e6
    log("fp32: %.3f ms/step, %.0f MLUPS, mean launch ms: %s"
        % (t32 / args.steps, mlups32, {k: round(v, 4) for k, v in k32.items() if v}))

    if not args.skip_fp64:
        extra.update({"fp64_mlups": round(mlups64, 1),
                      "fp32_over_fp64": round(mlups32 / mlups64, 3),
                      "fp64_kernel_set": "fused strip-march kernels (one-kernel step for "
                                         "npc=2), reference expression order",
                      "fp64_launch_ms": {k: round(v, 4) for k, v in k64.items() if v},
                      "fp64_steps": {"warmup": w64, "timed": s64}})
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
    kms = k32[dom]
    alg_bytes = COMPULSORY[dom] * 4 * N
    achieved = alg_bytes / (kms / 1e3)
    sweeps = (1 + npc) if dom == "step" else SWEEPS[dom]
    per_sweep_bytes = sweeps * SWEEP_VALUES * 4 * N
    prof = {}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("size") == L and t.get("npc") == npc:
            prof = t.get("kernels", {}).get(dom, {})
    traffic = prof.get("bytes_per_launch")
    valu_instr = prof.get("valu_instructions")
    roofline = {"bound": "hbm", "achieved": round(achieved / 1e9, 1),
                "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK, 4),
                "traffic": traffic,
                "kernel": KERNEL[dom], "kernel_ms": round(kms, 4),
                "alg_bytes_per_launch": alg_bytes,
                "alg_bytes_per_node": COMPULSORY[dom] * 4,
                "traffic_GBps": (round(traffic / (kms / 1e3) / 1e9, 1) if traffic else None),
                "traffic_frac": (round(traffic / (kms / 1e3) / HBM_PEAK, 4) if traffic else None),
                "per_sweep_equivalent": {
                    "bytes_per_node": sweeps * SWEEP_VALUES * 4, "sweeps": sweeps,
                    "GBps": round(per_sweep_bytes / (kms / 1e3) / 1e9, 1),
                    "note": "the step's data flow counted per sweep, including the predictor's "
                            "half-step state that the fused kernel keeps on chip"},
                "launch_ms": {k: round(v, 4) for k, v in k32.items() if v}}
    assert roofline["frac"] <= 1.0, roofline
    # the practical streaming rate of this GPU, measured now: 1 GiB
    # device-to-device copies (read + write), against the launch's counted
    # traffic rate
    copy = calibrate(quiet=True)
    roofline["device_copy_GBps"] = round(copy, 1)
    if traffic:
        roofline["traffic_over_device_copy"] = round(roofline["traffic_GBps"] /
                                                     roofline["device_copy_GBps"], 3)
    if valu_instr:
        issue_ms = valu_instr / VALU_PEAK_INSTR_PER_S * 1e3
        roofline["valu"] = {"instructions_per_launch": valu_instr,
                            "issue_ms_at_peak_clock": round(issue_ms, 4),
                            "frac": round(issue_ms / kms, 3),
                            "note": "SQ_INSTS_VALU x 4 cycles / (256 CUs x 4 SIMDs x 2.4 GHz) "
                                    "over the measured launch time; the step is near both "
                                    "this and the device copy rate"}

    if not args.skip_surrogate:
        log("surrogate %dx%d: training on 512^2 frames, %d timed steps" % (L, L, args.surrogate_steps))
        extra["surrogate"] = surrogate_leg(lyo, L, args.surrogate_steps,
                                           train_iters=args.surrogate_train,
                                           warmup=args.warmup)
        log("surrogate: %s" % extra["surrogate"])

    cpu = None
    if not args.skip_cpu:
        info = cpu_info()
        mt = args.cpu_threads or min(16, info["cpus_available"])
        cw, cs = args.cpu_warmup, args.cpu_steps
        L1 = min(args.cpu_size_1t, L)
        log("cpu baseline: oracle %dx%d npc=%d, %d threads, %d + %d steps"
            % (L, L, npc, mt, cw, cs))
        vm = cpu_time("lyo", L, cs, cw, mt, npc=npc)
        log("cpu: %.2f MLUPS (%d threads)" % (vm, mt))
        log("cpu baseline: oracle %dx%d npc=%d, 1 thread, %d + %d steps" % (L1, L1, npc, cw, cs))
        v1 = cpu_time("lyo", L1, cs, cw, 0, npc=npc)
        log("cpu: %.2f MLUPS (1 thread)" % v1)
        cpu = {"value": round(vm, 2), "unit": "MLUPS", "cores": mt, "kind": "port",
               "sample": "%dx%d FP64 npc=%d, %d warm-up + %d timed steps, OpenMP %d threads "
                         "(OMP_PLACES=cores, OMP_PROC_BIND=close)" % (L, L, npc, cw, cs, mt),
               "single_thread": {"value": round(v1, 2), "unit": "MLUPS", "cores": 1,
                                 "sample": "%dx%d FP64 npc=%d, %d warm-up + %d timed steps, "
                                           "serial (the reference's default)" % (L1, L1, npc,
                                                                                   cw, cs)},
               "build": "gcc -O3 -fopenmp -ffp-contract=off (oracle/Makefile)",
               "host": dict(info, cores_note="a GPU box allots 16 host CPUs per GPU"),
               "gpu_over_cpu": round(mlups32 / vm, 1)}
        # the reference's CPU path beside the walls and go-or-grow legs
        if "fp64_walls_bc4_mlups" in extra:
            vw = cpu_time("lyo", L, cs, cw, mt, npc=npc, bc=4)
            extra["cpu_walls_bc4"] = {"value": round(vw, 2), "unit": "MLUPS", "cores": mt,
                                      "kind": "port", "sample": "%dx%d FP64 bc=4 npc=%d, %d "
                                      "warm-up + %d timed steps, OpenMP" % (L, L, npc, cw, cs)}
            log("cpu walls: %.2f MLUPS" % vw)
        if "go_or_grow" in extra:
            vg = cpu_time("growth", 500, 20, 1, mt, relax_steps=200)
            extra["cpu_go_or_grow"] = {"value": round(vg, 2), "unit": "MLUPS", "cores": mt,
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
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
