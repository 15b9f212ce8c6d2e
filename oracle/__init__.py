"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle restates the reference CPU loop (seismicCPML_3d_isotropic_simpelmodel.cpp,
lines 142-622) in plain C; see ela3d_oracle.c.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module, and only as the checker or
the timed CPU baseline -- never as a product code path.

Configs are duck-typed: any object with the attribute names of
``ela3d_amd.Config`` works.  Arrays are numpy, x-fastest, shape (nz, ny, nx).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("ELA3D_ORACLE_LIB") or os.path.join(_HERE, "libela3d_oracle.so")

FIELD_NAMES = ["vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz"]
MEM_NAMES = [
    "dvx_dx", "dvy_dy", "dvz_dz", "dvy_dx", "dvx_dy", "dvz_dx", "dvx_dz", "dvz_dy", "dvy_dz",
    "dsxx_dx", "dsxy_dy", "dsxz_dz", "dsxy_dx", "dsyy_dy", "dsyz_dz", "dsxz_dx", "dsyz_dy",
    "dszz_dz",
]
# axis (0 x, 1 y, 2 z) and grid (0 full, 1 half) of the table each memory variable uses
MEM_TABLE = [
    (0, 1), (1, 0), (2, 0), (0, 0), (1, 1), (0, 0), (2, 1), (1, 1), (2, 1),
    (0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 0), (0, 1), (1, 0), (2, 1),
]

COMPAT_XHALF_A = 1
COMPAT_ZHALF_MIN = 2
COMPAT_GHOST = 4
COMPAT_REFERENCE = 7


class OrcCfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("nstep", C.c_int),
        ("npml", C.c_int),
        ("npower", C.c_double), ("k_max", C.c_double), ("alpha_max", C.c_double),
        ("rcoef", C.c_double), ("vpml", C.c_double),
        ("f0", C.c_double), ("t0", C.c_double), ("factor", C.c_double),
        ("angle_deg", C.c_double),
        ("src_kind", C.c_int), ("wavelet", C.c_int),
        ("isrc", C.c_int), ("jsrc", C.c_int), ("ksrc", C.c_int),
        ("nrec", C.c_int),
        ("rec", C.POINTER(C.c_int)),
        ("sample_every", C.c_int),
        ("energy_every", C.c_int),
        ("compat", C.c_uint),
        ("dirichlet", C.c_int),
        ("free_surface", C.c_int),
        ("threads", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"oracle library missing at {_LIB_PATH}; run __graft_entry__.build() "
                "or `make -C oracle`")
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_tables.argtypes = [C.POINTER(OrcCfg), C.c_int, dp]
        L.orc_wavelet.argtypes = [C.POINTER(OrcCfg), C.c_int, dp]
        for sfx, ct in (("f32", C.c_float), ("f64", C.c_double)):
            p = C.POINTER(ct)
            pp = C.POINTER(p)
            getattr(L, f"orc_stress_{sfx}").argtypes = [C.POINTER(OrcCfg), p, p, p, pp, pp]
            getattr(L, f"orc_velocity_{sfx}").argtypes = [C.POINTER(OrcCfg), p, pp, pp]
            getattr(L, f"orc_energy_{sfx}").argtypes = [C.POINTER(OrcCfg), p, p, p, pp, dp]
            getattr(L, f"orc_run_{sfx}").argtypes = [
                C.POINTER(OrcCfg), p, p, p, C.c_int, C.c_int, pp, pp, dp, dp]
            getattr(L, f"orc_run_{sfx}").restype = C.c_int
        _lib = L
    return _lib


def make_cfg(cfg, threads: int = 0):
    """Build an OrcCfg from a product-style config.  Returns (OrcCfg, keepalive)."""
    rec = np.ascontiguousarray(np.asarray(cfg.receivers, dtype=np.int32).reshape(-1, 3))
    c = OrcCfg()
    c.nx, c.ny, c.nz = int(cfg.nx), int(cfg.ny), int(cfg.nz)
    c.dx, c.dy, c.dz, c.dt = float(cfg.dx), float(cfg.dy), float(cfg.dz), float(cfg.dt)
    c.nstep = int(cfg.nstep)
    c.npml = int(cfg.npoints_pml)
    c.npower, c.k_max = float(cfg.npower), float(cfg.k_max_pml)
    c.alpha_max, c.rcoef, c.vpml = float(cfg.alpha_max_pml), float(cfg.rcoef), float(cfg.vpml)
    c.f0, c.t0, c.factor = float(cfg.f0), float(cfg.t0_effective), float(cfg.factor)
    c.angle_deg = float(cfg.angle_force)
    c.src_kind, c.wavelet = int(cfg.src_kind), int(cfg.wavelet)
    c.isrc, c.jsrc, c.ksrc = int(cfg.isrc), int(cfg.jsrc), int(cfg.ksrc)
    c.nrec = int(rec.shape[0])
    c.rec = rec.ctypes.data_as(C.POINTER(C.c_int)) if rec.size else C.POINTER(C.c_int)()
    c.sample_every = int(cfg.sample_every)
    c.energy_every = int(cfg.energy_every)
    c.compat = int(cfg.compat_flags)
    c.dirichlet = 1 if cfg.dirichlet else 0
    c.free_surface = 1 if getattr(cfg, "free_surface", False) else 0
    c.threads = int(threads)
    return c, rec


def tables(cfg, axis: int) -> dict:
    c, keep = make_cfg(cfg)
    n = (c.nx, c.ny, c.nz)[axis]
    out = np.zeros(10 * n, dtype=np.float64)
    rc = lib().orc_tables(C.byref(c), axis, out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise ValueError("orc_tables failed")
    names = ["d", "K", "alpha", "a", "b", "d_half", "K_half", "alpha_half", "a_half", "b_half"]
    return {nm: out[i * n:(i + 1) * n].copy() for i, nm in enumerate(names)}


def wavelet(cfg, n: int) -> np.ndarray:
    c, keep = make_cfg(cfg)
    out = np.zeros(n, dtype=np.float64)
    if lib().orc_wavelet(C.byref(c), n, out.ctypes.data_as(C.POINTER(C.c_double))):
        raise ValueError("orc_wavelet failed")
    return out


class State:
    """Full-grid oracle state: 9 fields + 18 memory variables."""

    def __init__(self, cfg, dtype=np.float64):
        shape = (int(cfg.nz), int(cfg.ny), int(cfg.nx))
        self.dtype = np.dtype(dtype)
        self.f = [np.zeros(shape, dtype=self.dtype) for _ in range(9)]
        self.m = [np.zeros(shape, dtype=self.dtype) for _ in range(18)]

    def _pp(self, arrs):
        ct = C.c_float if self.dtype == np.float32 else C.c_double
        p = C.POINTER(ct)
        a = (p * len(arrs))(*[x.ctypes.data_as(p) for x in arrs])
        return a

    def field(self, fid: int) -> np.ndarray:
        return self.f[fid] if fid < 9 else self.m[fid - 9]


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _model_arrays(model, dtype):
    return [np.ascontiguousarray(np.asarray(a), dtype=dtype) for a in model]


def _ptr(a):
    ct = C.c_float if a.dtype == np.float32 else C.c_double
    return a.ctypes.data_as(C.POINTER(ct))


def stress(cfg, model, st: State, threads: int = 0):
    c, keep = make_cfg(cfg, threads)
    cp, cs, rho = _model_arrays(model, st.dtype)
    getattr(lib(), f"orc_stress_{_sfx(st.dtype)}")(
        C.byref(c), _ptr(cp), _ptr(cs), _ptr(rho), st._pp(st.f), st._pp(st.m))


def velocity(cfg, model, st: State, threads: int = 0):
    c, keep = make_cfg(cfg, threads)
    rho = _model_arrays(model, st.dtype)[2]
    getattr(lib(), f"orc_velocity_{_sfx(st.dtype)}")(
        C.byref(c), _ptr(rho), st._pp(st.f), st._pp(st.m))


def energy(cfg, model, st: State) -> np.ndarray:
    c, keep = make_cfg(cfg)
    cp, cs, rho = _model_arrays(model, st.dtype)
    out = np.zeros(2, dtype=np.float64)
    getattr(lib(), f"orc_energy_{_sfx(st.dtype)}")(
        C.byref(c), _ptr(cp), _ptr(cs), _ptr(rho), st._pp(st.f),
        out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def run(cfg, model, nsteps=None, st: State | None = None, it0: int = 0, dtype=np.float64,
        threads: int = 0):
    """Run the oracle time loop; returns (seismograms [ns][nrec][3], energy [ne][2], state)."""
    nsteps = int(cfg.nstep) if nsteps is None else int(nsteps)
    if st is None:
        st = State(cfg, dtype)
    c, keep = make_cfg(cfg, threads)
    nrec = c.nrec
    last = it0 + nsteps
    ns = last // cfg.sample_every if cfg.sample_every > 0 else 0
    ne = last // cfg.energy_every if cfg.energy_every > 0 else 0
    seis = np.zeros((max(ns, 0), nrec, 3), dtype=np.float64)
    en = np.zeros((max(ne, 0), 2), dtype=np.float64)
    cp, cs, rho = _model_arrays(model, st.dtype)
    dp = C.POINTER(C.c_double)
    rc = getattr(lib(), f"orc_run_{_sfx(st.dtype)}")(
        C.byref(c), _ptr(cp), _ptr(cs), _ptr(rho), int(it0), int(nsteps),
        st._pp(st.f), st._pp(st.m),
        seis.ctypes.data_as(dp) if seis.size else dp(),
        en.ctypes.data_as(dp) if en.size else dp())
    if rc:
        raise ValueError("orc_run failed")
    return seis, en, st
