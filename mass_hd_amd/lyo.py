"""Python host binding of the C ABI in include/mass_lyo.h.

Mirrors the reference's Model interface for the lyotropic model
(src/models.hpp:30-116, src/models/lyotropic.{hpp,cpp}): ``Lyotropic``
takes the runcard options of ``Lyotropic::GetOptions``
(src/models/lyotropic.cpp:551-602), ``configure`` stands in for ``Configure``,
``step`` for ``Step``, ``runtime_checks`` for ``RuntimeChecks`` and
``frame`` for ``serialize_frame``.  Every call goes to the HIP library; there
is no CPU path, and constructing a model without the built library or without
a gfx950 device raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmass_lyo.so")

FP64, FP32 = 0, 1

CONFIGS = {
    "circle": 0, "square": 1, "stripe": 2, "half": 3, "random": 4,
    "boxatwall": 5, "cuttingatwall": 6, "circleatwall": 7, "wettedwall": 8,
}

FIELDS = {
    "ff": 0, "QQxx": 1, "QQyx": 2, "phi": 3, "n": 4, "ux": 5, "uy": 6,
    "sigmaXX": 7, "sigmaXY": 8, "sigmaYX": 9, "sigmaYY": 10,
    "sigma_bulk": 11, "sigma_elastic_xx": 12, "sigma_elastic_yx": 13,
    "sigma_phase_field_xx": 14, "sigma_phase_field_yx": 15,
    "sigma_active_xx": 16, "sigma_active_yx": 17, "MU": 18, "HHxx": 19, "HHyx": 20,
    "m": 21, "chi": 22, "division_mask": 23, "death_mask": 24,
}

# go-or-grow parameters (mass_lyo_growth; go-or-grow.hpp:10-28 defaults)
GROWTH_DOUBLES = ["B", "Snem", "Dchi", "Achi", "Ochi", "phiswitch", "chi0", "chi_noise",
                  "chi_length", "relax_dt", "alpha", "beta", "division_rate", "division_time",
                  "division_radius", "death_rate", "death_time", "death_radius", "phi_critical"]
GROWTH_DEFAULTS = dict(division_only=0, chi_config="noise", growTogether=0, relax_phi=0,
                       relax_Q=0, relax_steps=0,
                       B=0.0, Snem=1.0, Dchi=0.0, Achi=0.0, Ochi=0.0, phiswitch=0.0, chi0=0.0,
                       chi_noise=0.0, chi_length=0.0, relax_dt=1.0)


class Growth(ctypes.Structure):
    _fields_ = ([("chi_config", ctypes.c_int32), ("growTogether", ctypes.c_int32),
                 ("relax_phi", ctypes.c_int32), ("relax_Q", ctypes.c_int32),
                 ("relax_steps", ctypes.c_uint32), ("division_only", ctypes.c_int32)]
                + [(k, ctypes.c_double) for k in GROWTH_DOUBLES])

# the frame of Lyotropic::serialize_frame (src/models/lyotropic.hpp:121-138)
FRAME_FIELDS = ["ff", "QQxx", "QQyx", "phi", "sigmaXX", "sigmaXY", "sigmaYX", "sigmaYY",
                "sigma_bulk", "sigma_elastic_xx", "sigma_elastic_yx",
                "sigma_phase_field_xx", "sigma_phase_field_yx", "sigma_active_xx",
                "sigma_active_yx"]

PHYS_KEYS = ["GammaQ", "GammaP", "xi", "tauNem", "tauIso", "rho", "friction",
             "AA", "CC", "LL", "KK", "zeta", "zetaI", "conc"]


class Params(ctypes.Structure):
    _fields_ = ([("LX", ctypes.c_int32), ("LY", ctypes.c_int32), ("npc", ctypes.c_int32),
                 ("precision", ctypes.c_int32), ("conserve_phi", ctypes.c_int32),
                 ("plane_pad", ctypes.c_int32), ("bc", ctypes.c_int32)]
                + [(k, ctypes.c_double) for k in PHYS_KEYS])


class Init(ctypes.Structure):
    _fields_ = [("config", ctypes.c_int32), ("seed", ctypes.c_uint32),
                ("level", ctypes.c_double), ("radius", ctypes.c_double),
                ("angle_deg", ctypes.c_double), ("noise", ctypes.c_double),
                ("qamp", ctypes.c_double), ("phinoise", ctypes.c_double)]


HIST_BINS = 64


class Diag(ctypes.Structure):
    _fields_ = [("u_rms", ctypes.c_double), ("u_mean", ctypes.c_double),
                ("phi_mean", ctypes.c_double), ("phi_min", ctypes.c_double),
                ("phi_max", ctypes.c_double), ("defect_density", ctypes.c_double),
                ("defects_pos", ctypes.c_int64), ("defects_neg", ctypes.c_int64),
                ("phi_hist", ctypes.c_uint64 * (HIST_BINS + 2))]


class MassLyoError(RuntimeError):
    """A non-zero status of the C ABI; ``code`` is the mass_lyo_status."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def load_library():
    """Loads libmass_lyo.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmass_lyo.so is not built (%s); run __graft_entry__.build()"
                           % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    h = ctypes.c_void_p
    lib.mass_lyo_create.argtypes = [P(Params), P(h)]
    lib.mass_lyo_destroy.argtypes = [h]
    lib.mass_lyo_destroy.restype = None
    lib.mass_lyo_last_error.argtypes = [h]
    lib.mass_lyo_last_error.restype = ctypes.c_char_p
    lib.mass_lyo_configure.argtypes = [h, P(Init), dp, dp]
    lib.mass_lyo_set_state.argtypes = [h, dp, dp, dp, dp]
    lib.mass_lyo_step.argtypes = [h, ctypes.c_int32]
    lib.mass_lyo_get_field.argtypes = [h, ctypes.c_int32, dp]
    lib.mass_lyo_sums.argtypes = [h, dp, dp]
    lib.mass_lyo_synchronize.argtypes = [h]
    lib.mass_lyo_time_steps.argtypes = [h, ctypes.c_int32, dp, dp]
    lib.mass_lyo_set_launch.argtypes = [h, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    lib.mass_lyo_set_schedule.argtypes = [h, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_int32]
    lib.mass_lyo_set_loop_kernels.argtypes = [h, ctypes.c_int32]
    lib.mass_lyo_set_growth.argtypes = [h, ctypes.POINTER(Growth)]
    lib.mass_lyo_checkpoint_size.argtypes = [h, ctypes.POINTER(ctypes.c_uint64)]
    lib.mass_lyo_checkpoint_save.argtypes = [h, ctypes.c_void_p, ctypes.c_uint64]
    lib.mass_lyo_checkpoint_load.argtypes = [h, ctypes.c_void_p, ctypes.c_uint64]
    lib.mass_lyo_abi_version.argtypes = []
    lib.mass_lyo_diagnostics.argtypes = [h, P(Diag), dp]
    i32 = ctypes.c_int32
    lib.mass_lyo_batch_create.argtypes = [P(Params), i32, P(h)]
    lib.mass_lyo_batch_destroy.argtypes = [h]
    lib.mass_lyo_batch_destroy.restype = None
    lib.mass_lyo_batch_last_error.argtypes = [h]
    lib.mass_lyo_batch_last_error.restype = ctypes.c_char_p
    lib.mass_lyo_batch_size.argtypes = [h]
    lib.mass_lyo_batch_configure.argtypes = [h, i32, P(Init), dp, dp]
    lib.mass_lyo_batch_set_state.argtypes = [h, i32, dp, dp, dp, dp]
    lib.mass_lyo_batch_step.argtypes = [h, i32]
    lib.mass_lyo_batch_get_field.argtypes = [h, i32, i32, dp]
    lib.mass_lyo_batch_sums.argtypes = [h, i32, dp, dp]
    lib.mass_lyo_batch_diagnostics.argtypes = [h, i32, P(Diag), dp]
    lib.mass_lyo_batch_set_launch.argtypes = [h, i32, i32]
    lib.mass_lyo_batch_autotune.argtypes = [h, P(i32), P(i32)]
    lib.mass_lyo_batch_time_steps.argtypes = [h, i32, dp]
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Lyotropic:
    """One lyotropic run on the GPU (the reference's ``Lyotropic`` model).

    bc: 0 periodic, 1 free-slip channel, 2 no-slip channel, 3 free-slip box,
    4 no-slip box (walled grids: FP64 only)."""

    def __init__(self, LX, LY, npc=1, precision=FP64, plane_pad=0, bc=0, **phys):
        self.lib = load_library()
        if isinstance(precision, str):
            precision = {"fp64": FP64, "fp32": FP32}[precision]
        missing = [k for k in PHYS_KEYS if k not in phys]
        if missing:
            raise ValueError("missing model parameters: %s" % ", ".join(missing))
        unknown = set(phys) - set(PHYS_KEYS)
        if unknown:
            raise ValueError("unknown model parameters: %s" % ", ".join(sorted(unknown)))
        self.LX, self.LY, self.N = int(LX), int(LY), int(LX) * int(LY)
        self.npc, self.precision, self.bc = int(npc), precision, int(bc)
        self.phys = dict(phys)
        p = Params(LX=self.LX, LY=self.LY, npc=self.npc, precision=precision,
                   conserve_phi=0, plane_pad=int(plane_pad), bc=self.bc,
                   **{k: float(phys[k]) for k in PHYS_KEYS})
        h = ctypes.c_void_p()
        rc = self.lib.mass_lyo_create(ctypes.byref(p), ctypes.byref(h))
        if rc != 0:
            raise MassLyoError(rc, self.lib.mass_lyo_last_error(None).decode())
        self.h = h
        self.ftot = 0.0
        self.ptot = 0.0

    def _check(self, rc):
        if rc != 0:
            raise MassLyoError(rc, self.lib.mass_lyo_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mass_lyo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- Lyotropic::Configure
    def configure(self, config="random", seed=1, level=0.0, radius=0.0, angle=0.0, noise=1.0,
                  qamp=1.0, phinoise=0.0):
        cfg = CONFIGS[config] if isinstance(config, str) else int(config)
        ini = Init(config=cfg, seed=seed, level=level, radius=radius, angle_deg=angle,
                   noise=noise, qamp=qamp, phinoise=phinoise)
        ft, pt = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.mass_lyo_configure(self.h, ctypes.byref(ini), ctypes.byref(ft),
                                                ctypes.byref(pt)))
        self.ftot, self.ptot = ft.value, pt.value
        return self.ftot, self.ptot

    def set_state(self, ff, Qxx, Qyx, phi):
        ff = np.ascontiguousarray(ff, dtype=np.float64).reshape(self.N * 9)
        a = [np.ascontiguousarray(x, dtype=np.float64).reshape(self.N) for x in (Qxx, Qyx, phi)]
        self._check(self.lib.mass_lyo_set_state(self.h, _dptr(ff), _dptr(a[0]), _dptr(a[1]),
                                                _dptr(a[2])))
        self.ftot = float(ff.sum())
        self.ptot = float(a[2].sum())

    # ---- Lyotropic::Step
    def step(self, nsteps=1):
        self._check(self.lib.mass_lyo_step(self.h, int(nsteps)))

    def synchronize(self):
        self._check(self.lib.mass_lyo_synchronize(self.h))

    def get(self, name):
        fid = FIELDS[name] if isinstance(name, str) else int(name)
        n = self.N * (9 if fid == 0 else 1)
        out = np.empty(n, dtype=np.float64)
        self._check(self.lib.mass_lyo_get_field(self.h, fid, _dptr(out)))
        return out.reshape(self.N, 9) if fid == 0 else out

    # ---- Lyotropic::serialize_frame
    def frame(self):
        return {k: self.get(k) for k in FRAME_FIELDS}

    def sums(self):
        f, p = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.mass_lyo_sums(self.h, ctypes.byref(f), ctypes.byref(p)))
        return f.value, p.value

    # ---- Lyotropic::RuntimeChecks (src/models/lyotropic.cpp:527-549)
    def runtime_checks(self):
        fcheck, pcheck = self.sums()
        if abs(self.ftot - fcheck) > 1:
            raise MassLyoError(4, "f is not conserved (%g/%g)" % (self.ftot, fcheck))
        return fcheck, pcheck

    def diagnostics(self, charge=False):
        """u_rms, phi statistics/histogram and +-1/2 defect counts of the
        current state; with charge=True also the winding-number field."""
        d = Diag()
        w = np.empty(self.N) if charge else None
        self._check(self.lib.mass_lyo_diagnostics(self.h, ctypes.byref(d),
                                                  _dptr(w) if charge else None))
        out = {k: getattr(d, k) for k, _ in Diag._fields_ if k != "phi_hist"}
        out["phi_hist"] = np.array(d.phi_hist[:], dtype=np.int64)
        if charge:
            out["charge"] = w
        return out

    def time_steps(self, nsteps, per_sweep=True):
        tot = ctypes.c_double()
        sw = (ctypes.c_double * 4)()
        self._check(self.lib.mass_lyo_time_steps(self.h, int(nsteps), ctypes.byref(tot),
                                                 sw if per_sweep else None))
        return tot.value, (sw[0], sw[1], sw[2], sw[3])

    def set_launch(self, chunk_rows=0, waves_per_block=0, xcd_order=-1):
        self._check(self.lib.mass_lyo_set_launch(self.h, int(chunk_rows), int(waves_per_block),
                                                 int(xcd_order)))

    def set_schedule(self, fuse_last_pair=-1, pair_chunk_rows=0, alternate=-1,
                     one_kernel_step=-1, step_chunk_rows=0):
        """fuse_last_pair 1: the last two correctors run as one kernel
        (default), 0: separate sweeps; pair_chunk_rows: rows per wave chunk of
        the fused kernel (0 keeps); alternate 1: neighbouring chunks march in
        opposite directions (default), 0: all upwards; one_kernel_step 1:
        npc=1 or 2 runs predictor and correctors in one kernel (default), 0:
        predictor + fused pair or corrector sweep, 2 (npc = 2): its
        block-cooperative form (waves trade edge values through LDS);
        step_chunk_rows: its rows per chunk.  -1
        keeps a setting.  Results are bitwise identical either way."""
        self._check(self.lib.mass_lyo_set_schedule(self.h, int(fuse_last_pair),
                                                   int(pair_chunk_rows), int(alternate),
                                                   int(one_kernel_step), int(step_chunk_rows)))

    def checkpoint(self):
        """binary image of the run's state (header + raw device planes);
        restore() on a run of the same grid, precision and rho continues
        bitwise as this run would"""
        n = ctypes.c_uint64()
        self._check(self.lib.mass_lyo_checkpoint_size(self.h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._check(self.lib.mass_lyo_checkpoint_save(self.h, buf, n.value))
        return buf.raw

    def restore(self, image):
        image = bytes(image)
        self._check(self.lib.mass_lyo_checkpoint_load(
            self.h, ctypes.cast(ctypes.c_char_p(image), ctypes.c_void_p), len(image)))

    def set_growth(self, **kw):
        """Makes this a go-or-grow run (before configure; division_only=1: a
        lyotropic-with-division run): keys of
        GROWTH_DEFAULTS (chi_config "noise" or "front") and alpha, beta,
        division_rate/time/radius, death_rate/time/radius, phi_critical."""
        g = dict(GROWTH_DEFAULTS)
        g.update(kw)
        missing = [k for k in GROWTH_DOUBLES if k not in g]
        if missing:
            raise ValueError("missing go-or-grow parameters: %s" % ", ".join(missing))
        cfg = {"noise": 0, "front": 1}[g["chi_config"]]
        s = Growth(chi_config=cfg, growTogether=int(g["growTogether"]),
                   relax_phi=int(g["relax_phi"]), relax_Q=int(g["relax_Q"]),
                   relax_steps=int(g["relax_steps"]),
                   division_only=int(g["division_only"]),
                   **{k: float(g[k]) for k in GROWTH_DOUBLES})
        self._check(self.lib.mass_lyo_set_growth(self.h, ctypes.byref(s)))

    def set_loop_kernels(self, on=True):
        """FP64: run each step as the reference's node loops, one kernel per
        loop (UpdateQuantities storing its 23 derived fields, UpdateFields,
        Move); False returns to the fused kernels (on walled runs the
        ghost-layer tile sweeps, on periodic go-or-grow runs their tile
        sweeps; walled go-or-grow runs refuse False).  Bitwise the same
        results."""
        self._check(self.lib.mass_lyo_set_loop_kernels(self.h, 1 if on else 0))


class Batch:
    """Independent runs advanced together (persistent batched sweeps + HIP
    graph).  ``runs`` is a list of dicts with LX, LY and the model parameters;
    npc and precision are shared."""

    def __init__(self, runs, npc=2, precision=FP32):
        self.lib = load_library()
        if isinstance(precision, str):
            precision = {"fp64": FP64, "fp32": FP32}[precision]
        arr = (Params * len(runs))()
        self.sizes = []
        for i, r in enumerate(runs):
            phys = {k: float(r[k]) for k in PHYS_KEYS}
            arr[i] = Params(LX=int(r["LX"]), LY=int(r["LY"]), npc=int(npc), precision=precision,
                            conserve_phi=0, plane_pad=0, bc=int(r.get("bc", 0)), **phys)
            self.sizes.append((int(r["LX"]), int(r["LY"])))
        h = ctypes.c_void_p()
        rc = self.lib.mass_lyo_batch_create(arr, len(runs), ctypes.byref(h))
        if rc != 0:
            raise MassLyoError(rc, self.lib.mass_lyo_batch_last_error(None).decode())
        self.h = h
        self.npc, self.precision = npc, precision

    def _check(self, rc):
        if rc != 0:
            raise MassLyoError(rc, self.lib.mass_lyo_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mass_lyo_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.sizes)

    def configure(self, run, config="random", seed=1, level=0.0, radius=0.0, angle=0.0,
                  noise=1.0, qamp=1.0, phinoise=0.0):
        cfg = CONFIGS[config] if isinstance(config, str) else int(config)
        ini = Init(config=cfg, seed=seed, level=level, radius=radius, angle_deg=angle,
                   noise=noise, qamp=qamp, phinoise=phinoise)
        ft, pt = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.mass_lyo_batch_configure(self.h, run, ctypes.byref(ini),
                                                      ctypes.byref(ft), ctypes.byref(pt)))
        return ft.value, pt.value

    def set_state(self, run, ff, Qxx, Qyx, phi):
        LX, LY = self.sizes[run]
        ff = np.ascontiguousarray(ff, dtype=np.float64).reshape(LX * LY * 9)
        a = [np.ascontiguousarray(x, dtype=np.float64).reshape(LX * LY) for x in (Qxx, Qyx, phi)]
        self._check(self.lib.mass_lyo_batch_set_state(self.h, run, _dptr(ff), _dptr(a[0]),
                                                      _dptr(a[1]), _dptr(a[2])))

    def step(self, nsteps=1):
        self._check(self.lib.mass_lyo_batch_step(self.h, int(nsteps)))

    def get(self, run, name):
        LX, LY = self.sizes[run]
        fid = FIELDS[name] if isinstance(name, str) else int(name)
        out = np.empty(LX * LY * (9 if fid == 0 else 1))
        self._check(self.lib.mass_lyo_batch_get_field(self.h, run, fid, _dptr(out)))
        return out.reshape(LX * LY, 9) if fid == 0 else out

    def sums(self, run):
        f, p = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.mass_lyo_batch_sums(self.h, run, ctypes.byref(f), ctypes.byref(p)))
        return f.value, p.value

    def diagnostics(self, run):
        d = Diag()
        self._check(self.lib.mass_lyo_batch_diagnostics(self.h, run, ctypes.byref(d), None))
        out = {k: getattr(d, k) for k, _ in Diag._fields_ if k != "phi_hist"}
        out["phi_hist"] = np.array(d.phi_hist[:], dtype=np.int64)
        return out

    def set_launch(self, chunk_rows=0, grid_waves=0):
        self._check(self.lib.mass_lyo_batch_set_launch(self.h, int(chunk_rows), int(grid_waves)))

    def autotune(self):
        c, g = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.lib.mass_lyo_batch_autotune(self.h, ctypes.byref(c), ctypes.byref(g)))
        return c.value, g.value

    def time_steps(self, nsteps):
        t = ctypes.c_double()
        self._check(self.lib.mass_lyo_batch_time_steps(self.h, int(nsteps), ctypes.byref(t)))
        return t.value


class SurrogateStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("surrogate_steps", ctypes.c_int64),
                ("plain_steps", ctypes.c_int64), ("judged", ctypes.c_int64),
                ("fallbacks", ctypes.c_int64), ("last_error", ctypes.c_double),
                ("max_error", ctypes.c_double)]


def _net_api(lib):
    if getattr(lib, "_net_api", False):
        return lib
    P = ctypes.POINTER
    h = ctypes.c_void_p
    fp = P(ctypes.c_float)
    dp = P(ctypes.c_double)
    i32 = ctypes.c_int32
    lib.mass_lyo_net_create.argtypes = [ctypes.c_uint32, P(h)]
    lib.mass_lyo_net_destroy.argtypes = [h]
    lib.mass_lyo_net_destroy.restype = None
    lib.mass_lyo_net_last_error.argtypes = [h]
    lib.mass_lyo_net_last_error.restype = ctypes.c_char_p
    lib.mass_lyo_net_num_params.argtypes = [h]
    lib.mass_lyo_net_get_params.argtypes = [h, fp]
    lib.mass_lyo_net_set_params.argtypes = [h, fp]
    lib.mass_lyo_net_normalization.argtypes = [h, fp]
    lib.mass_lyo_net_forward_host.argtypes = [h, fp, i32, i32, fp]
    lib.mass_lyo_net_predict_host.argtypes = [h, fp, i32, i32, fp]
    lib.mass_lyo_net_grad_host.argtypes = [h, fp, fp, i32, i32, dp, fp]
    lib.mass_lyo_net_collect.argtypes = [h, h]
    lib.mass_lyo_net_dataset_size.argtypes = [h]
    lib.mass_lyo_net_train.argtypes = [h, i32, ctypes.c_float, dp, dp]
    lib.mass_lyo_surrogate_enable.argtypes = [h, h, i32, ctypes.c_double]
    lib.mass_lyo_surrogate_disable.argtypes = [h]
    lib.mass_lyo_surrogate_get_stats.argtypes = [h, P(SurrogateStats)]
    lib.mass_lyo_surrogate_set_kernels.argtypes = [h, i32]
    lib.mass_lyo_surrogate_set_calibration.argtypes = [h, i32]
    lib.mass_lyo_surrogate_get_calibration.argtypes = [h, dp]
    lib._net_api = True
    return lib


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


NET_CHANNELS = [8, 32, 32, 32, 8]


class Net:
    """The surrogate corrector network (include/mass_lyo.h, mass_lyo_net_*)."""

    def __init__(self, seed=0):
        self.lib = _net_api(load_library())
        h = ctypes.c_void_p()
        rc = self.lib.mass_lyo_net_create(seed, ctypes.byref(h))
        if rc != 0:
            raise MassLyoError(rc, self.lib.mass_lyo_net_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise MassLyoError(rc, self.lib.mass_lyo_net_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mass_lyo_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self):
        return self.lib.mass_lyo_net_num_params(self.h)

    def get_params(self):
        out = np.empty(self.num_params, dtype=np.float32)
        self._check(self.lib.mass_lyo_net_get_params(self.h, _fptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.lib.mass_lyo_net_set_params(self.h, _fptr(p)))

    def layers(self, flat=None):
        """[(W[9, cout, cin], b[cout])] views of a flat parameter vector."""
        flat = self.get_params() if flat is None else flat
        out, o = [], 0
        for l in range(4):
            ci, co = NET_CHANNELS[l], NET_CHANNELS[l + 1]
            W = flat[o:o + 9 * co * ci].reshape(9, co, ci)
            o += 9 * co * ci
            b = flat[o:o + co]
            o += co
            out.append((W, b))
        return out

    def normalization(self):
        out = np.empty(32, dtype=np.float32)
        self._check(self.lib.mass_lyo_net_normalization(self.h, _fptr(out)))
        return out

    def forward(self, x):
        LX, LY, _ = x.shape
        xi = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((LX, LY, 8), dtype=np.float32)
        self._check(self.lib.mass_lyo_net_forward_host(self.h, _fptr(xi), LX, LY, _fptr(y)))
        return y

    def predict(self, x):
        """surrogate inference on features in physical units (normalised with
        the network's statistics, output scaled to increments)"""
        LX, LY, _ = x.shape
        xi = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((LX, LY, 8), dtype=np.float32)
        self._check(self.lib.mass_lyo_net_predict_host(self.h, _fptr(xi), LX, LY, _fptr(y)))
        return y

    def grad(self, x, y):
        LX, LY, _ = x.shape
        xi = np.ascontiguousarray(x, dtype=np.float32)
        yi = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(self.num_params, dtype=np.float32)
        L = ctypes.c_double()
        self._check(self.lib.mass_lyo_net_grad_host(self.h, _fptr(xi), _fptr(yi), LX, LY,
                                                     ctypes.byref(L), _fptr(g)))
        return L.value, g

    def collect(self, model):
        self._check(self.lib.mass_lyo_net_collect(self.h, model.h))

    @property
    def dataset_size(self):
        return self.lib.mass_lyo_net_dataset_size(self.h)

    def train(self, iters, lr=1e-3):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.mass_lyo_net_train(self.h, int(iters), float(lr), ctypes.byref(a),
                                                ctypes.byref(b)))
        return a.value, b.value


def enable_surrogate(model, net, judge_every=50, tol=1e-3):
    lib = _net_api(model.lib)
    model._check(lib.mass_lyo_surrogate_enable(model.h, net.h, int(judge_every), float(tol)))


def set_surrogate_kernels(model, fused):
    """1: the network part of surrogate steps in one kernel (the FP32
    default); 0: separate feature, convolution and increment kernels."""
    lib = _net_api(model.lib)
    model._check(lib.mass_lyo_surrogate_set_kernels(model.h, int(fused)))


def set_surrogate_calibration(model, on):
    """Judged steps also fit per-output factors of the network's increments
    (least squares against the plain corrector's); 1 on, 0 off (resets)."""
    lib = _net_api(model.lib)
    model._check(lib.mass_lyo_surrogate_set_calibration(model.h, int(on)))


def surrogate_calibration(model):
    lib = _net_api(model.lib)
    a = (ctypes.c_double * 5)()
    model._check(lib.mass_lyo_surrogate_get_calibration(model.h, a))
    return list(a)


def disable_surrogate(model):
    lib = _net_api(model.lib)
    model._check(lib.mass_lyo_surrogate_disable(model.h))


def surrogate_stats(model):
    lib = _net_api(model.lib)
    s = SurrogateStats()
    model._check(lib.mass_lyo_surrogate_get_stats(model.h, ctypes.byref(s)))
    return {k: getattr(s, k) for k, _ in SurrogateStats._fields_}


class RunRef:
    """Borrowed handle of one member of a Batch (reading, diagnostics,
    collecting training frames)."""

    def __init__(self, batch, i):
        lib = batch.lib
        lib.mass_lyo_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        lib.mass_lyo_batch_run.restype = ctypes.c_void_p
        h = lib.mass_lyo_batch_run(batch.h, int(i))
        if not h:
            raise MassLyoError(1, "no run %d in the batch" % i)
        self.h = ctypes.c_void_p(h)
        self.lib = lib
        self.batch = batch  # keeps the batch alive
