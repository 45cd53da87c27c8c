"""ctypes binding of the CPU oracle (oracle/liblyo_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, always as the checker.  The product never
imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LYO_ORACLE_LIB points the tests at another build of the oracle (the
# planted-mistake checks build altered copies of lyo_oracle.c)
LIB_PATH = os.environ.get("LYO_ORACLE_LIB") or os.path.join(_HERE, "liblyo_oracle.so")

PHYS_KEYS = ["GammaQ", "GammaP", "xi", "tauNem", "tauIso", "rho", "friction",
             "AA", "CC", "LL", "KK", "zeta", "zetaI", "conc"]

FIELDS = {
    "ff": 0, "fn": 1, "QQxx": 2, "QQyx": 3, "phi": 4, "QNxx": 5, "QNyx": 6, "phn": 7,
    "n": 8, "ux": 9, "uy": 10, "ux_phi": 11, "uy_phi": 12, "HHxx": 13, "HHyx": 14, "MU": 15,
    "dxQQxx": 16, "dyQQxx": 17, "dxQQyx": 18, "dyQQyx": 19,
    "sigmaXX": 20, "sigmaYY": 21, "sigmaYX": 22, "sigmaXY": 23,
    "sigma_bulk": 24, "sigma_elastic_xx": 25, "sigma_elastic_yx": 26,
    "sigma_phase_field_xx": 27, "sigma_phase_field_yx": 28,
    "sigma_active_xx": 29, "sigma_active_yx": 30, "m": 31, "mN": 32, "chi": 33,
}

CONFIGS = {
    "circle": 0, "square": 1, "stripe": 2, "half": 3, "random": 4,
    "boxatwall": 5, "cuttingatwall": 6, "circleatwall": 7, "wettedwall": 8,
}


class _Params(ctypes.Structure):
    _fields_ = ([("LX", ctypes.c_int32), ("LY", ctypes.c_int32), ("npc", ctypes.c_int32),
                 ("nthreads", ctypes.c_int32), ("conserve_phi", ctypes.c_int32),
                 ("bc", ctypes.c_int32)] + [(k, ctypes.c_double) for k in PHYS_KEYS])


GROWTH_DOUBLES = ["B", "Snem", "Dchi", "Achi", "Ochi", "phiswitch", "chi0", "chi_noise",
                  "chi_length", "relax_dt", "alpha", "beta", "division_rate", "division_time",
                  "division_radius", "death_rate", "death_time", "death_radius", "phi_critical"]
# go-or-grow defaults (go-or-grow.hpp:10-28)
GROWTH_DEFAULTS = dict(division_only=0, chi_config="noise", growTogether=0, relax_phi=0,
                       relax_Q=0, relax_steps=0,
                       B=0.0, Snem=1.0, Dchi=0.0, Achi=0.0, Ochi=0.0, phiswitch=0.0, chi0=0.0,
                       chi_noise=0.0, chi_length=0.0, relax_dt=1.0)


class _Growth(ctypes.Structure):
    _fields_ = ([("chi_config", ctypes.c_int32), ("growTogether", ctypes.c_int32),
                 ("relax_phi", ctypes.c_int32), ("relax_Q", ctypes.c_int32),
                 ("relax_steps", ctypes.c_uint32), ("division_only", ctypes.c_int32)]
                + [(k, ctypes.c_double) for k in GROWTH_DOUBLES])


class _Init(ctypes.Structure):
    _fields_ = [("config", ctypes.c_int32), ("seed", ctypes.c_uint32),
                ("level", ctypes.c_double), ("radius", ctypes.c_double),
                ("angle_deg", ctypes.c_double), ("noise", ctypes.c_double),
                ("qamp", ctypes.c_double), ("phinoise", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        up = ctypes.POINTER(ctypes.c_uint32)
        L.lyo_oracle_create.argtypes = [ctypes.POINTER(_Params)]
        L.lyo_oracle_create.restype = vp
        L.lyo_oracle_destroy.argtypes = [vp]
        L.lyo_oracle_destroy.restype = None
        L.lyo_oracle_configure.argtypes = [vp, ctypes.POINTER(_Init)]
        L.lyo_oracle_set_state.argtypes = [vp, dp, dp, dp, dp]
        L.lyo_oracle_set_state.restype = None
        L.lyo_oracle_get.argtypes = [vp, ctypes.c_int, dp]
        L.lyo_oracle_set_field.argtypes = [vp, ctypes.c_int, dp]
        L.lyo_oracle_step.argtypes = [vp, ctypes.c_int]
        L.lyo_oracle_step.restype = None
        for f in ("lyo_oracle_update_quantities", "lyo_oracle_move"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = None
        L.lyo_oracle_update_fields.argtypes = [vp, ctypes.c_int]
        L.lyo_oracle_update_fields.restype = None
        L.lyo_oracle_sums.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.lyo_oracle_sums.restype = None
        L.lyo_oracle_neighbours.argtypes = [vp, up]
        L.lyo_oracle_neighbours.restype = None
        L.lyo_oracle_total_size.argtypes = [vp]
        L.lyo_oracle_total_size.restype = ctypes.c_uint32
        L.lyo_oracle_boundary.argtypes = [vp, ctypes.c_int]
        L.lyo_oracle_boundary.restype = None
        L.lyo_oracle_get_all.argtypes = [vp, ctypes.c_int, dp]
        L.lyo_oracle_set_growth.argtypes = [vp, ctypes.POINTER(_Growth)]
        L.lyo_oracle_masks.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p]
        L.lyo_oracle_masks.restype = None
        L.lyo_oracle_random_normal.argtypes = [ctypes.c_uint32, ctypes.c_int, dp]
        L.lyo_oracle_random_normal.restype = None
        L.lyo_oracle_mt19937.argtypes = [ctypes.c_uint32, ctypes.c_int, up]
        L.lyo_oracle_mt19937.restype = None
        L.lyo_oracle_random_real.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, dp]
        L.lyo_oracle_random_real.restype = None
        L.lyo_oracle_equilibrium.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        L.lyo_oracle_equilibrium.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """FP64 CPU restatement of the reference Lyotropic model."""

    def __init__(self, LX, LY, npc=1, nthreads=0, bc=0, **phys):
        """bc: 0 periodic, 1 free-slip channel, 2 no-slip channel, 3 free-slip
        box, 4 no-slip box (lyotropic.cpp:388-498)."""
        self.L = lib()
        self.LX, self.LY, self.N, self.bc = LX, LY, LX * LY, bc
        p = _Params(LX=LX, LY=LY, npc=npc, nthreads=nthreads, conserve_phi=0, bc=bc,
                    **{k: float(phys[k]) for k in PHYS_KEYS})
        self.o = self.L.lyo_oracle_create(ctypes.byref(p))
        if not self.o:
            raise ValueError("oracle: bad grid size or bc")
        self.T = int(self.L.lyo_oracle_total_size(self.o))

    def close(self):
        if self.o:
            self.L.lyo_oracle_destroy(self.o)
            self.o = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, config="random", seed=1, level=0.0, radius=0.0, angle=0.0, noise=1.0,
                  qamp=1.0, phinoise=0.0):
        cfg = CONFIGS[config] if isinstance(config, str) else int(config)
        ini = _Init(config=cfg, seed=seed, level=level, radius=radius, angle_deg=angle,
                    noise=noise, qamp=qamp, phinoise=phinoise)
        if self.L.lyo_oracle_configure(self.o, ctypes.byref(ini)) != 0:
            raise ValueError("oracle: unknown config")
        s = self.sums()
        return s["ftot"], s["ptot"]

    def set_state(self, ff, Qxx, Qyx, phi):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ff, Qxx, Qyx, phi)]
        self.L.lyo_oracle_set_state(self.o, *[_dp(x) for x in a])

    def get(self, name):
        fid = FIELDS[name]
        out = np.empty(self.N * (9 if fid < 2 else 1))
        self.L.lyo_oracle_get(self.o, fid, _dp(out))
        return out.reshape(self.N, 9) if fid < 2 else out

    def set_field(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.L.lyo_oracle_set_field(self.o, FIELDS[name], _dp(a))

    def step(self, n=1):
        self.L.lyo_oracle_step(self.o, n)

    def update_quantities(self):
        self.L.lyo_oracle_update_quantities(self.o)

    def update_fields(self, first):
        self.L.lyo_oracle_update_fields(self.o, 1 if first else 0)

    def move(self):
        self.L.lyo_oracle_move(self.o)

    def sums(self):
        v = [ctypes.c_double() for _ in range(6)]
        self.L.lyo_oracle_sums(self.o, *[ctypes.byref(x) for x in v])
        return dict(zip(["fsum", "phisum", "ftot", "ptot", "totalphi", "countphi"],
                        [x.value for x in v]))

    def set_growth(self, **kw):
        """Makes this a go-or-grow run (before configure; division_only=1: a
        lyotropic-with-division run).  Keys: those of
        GROWTH_DEFAULTS (chi_config "noise" or "front") and alpha, beta,
        division_rate/time/radius, death_rate/time/radius, phi_critical."""
        g = dict(GROWTH_DEFAULTS)
        g.update(kw)
        cfg = {"noise": 0, "front": 1}[g["chi_config"]]
        s = _Growth(chi_config=cfg, growTogether=int(g["growTogether"]),
                    relax_phi=int(g["relax_phi"]), relax_Q=int(g["relax_Q"]),
                    relax_steps=int(g["relax_steps"]), division_only=int(g["division_only"]),
                    **{k: float(g[k]) for k in GROWTH_DOUBLES})
        if self.L.lyo_oracle_set_growth(self.o, ctypes.byref(s)) != 0:
            raise ValueError("oracle: bad growth parameters")

    def masks(self):
        div = np.zeros(self.N, dtype=np.uint8)
        death = np.zeros(self.N, dtype=np.uint8)
        self.L.lyo_oracle_masks(self.o, div.ctypes.data, death.ctypes.data)
        return div, death

    def boundary(self, which):
        """0: BoundaryConditionsFields, 1: ...Fields2, 2: ...LB."""
        self.L.lyo_oracle_boundary(self.o, which)

    def get_all(self, name):
        """Whole grid with the ghost layer (domain, walls L, R, F, B, corners
        LF, RF, LB, RB)."""
        fid = FIELDS[name]
        out = np.empty(self.T * (9 if fid < 2 else 1))
        self.L.lyo_oracle_get_all(self.o, fid, _dp(out))
        return out.reshape(self.T, 9) if fid < 2 else out

    def neighbours(self):
        out = np.empty(self.T * 9, dtype=np.uint32)
        self.L.lyo_oracle_neighbours(self.o, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        return out.reshape(self.T, 9)


def mt19937(seed, n):
    out = np.empty(n, dtype=np.uint32)
    lib().lyo_oracle_mt19937(seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def random_real(seed, n, lo=0.0, hi=1.0):
    out = np.empty(n)
    lib().lyo_oracle_random_real(seed, n, lo, hi, _dp(out))
    return out


def random_normal(seed, n):
    out = np.empty(n)
    lib().lyo_oracle_random_normal(seed, n, _dp(out))
    return out


def equilibrium(vx, vy, nn):
    out = np.empty(9)
    lib().lyo_oracle_equilibrium(vx, vy, nn, _dp(out))
    return out
