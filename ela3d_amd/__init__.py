"""ela3d_amd -- MI355X-native SEISMIC_CPML 3-D isotropic elastic time loop.

Python mirror of the C ABI in include/ela3d.h.  The reference program
(seismicCPML_3d_isotropic_simpelmodel.cpp) exposes its loop only through the
literals of main(); ``Config`` carries those same parameters (NIMX/NIMY/NIMZ,
DELTAX/Y/Z, DELTAT, NSTEP, NPOINTS_PML, ...), ``Shot`` runs the loop on the GPU
and returns seismograms and the total energy.  Errors raise ``Ela3dError``
(the reference printed and exited, HOM:153-162).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field, replace
from typing import Optional

import numpy as np

from ._lib import (ABI_VERSION, ConfigStruct, Ela3dError, InfoStruct, KernelStats, LayeredSpec,
                   check, lib)

__all__ = [
    "Config", "Shot", "Ela3dError", "homogeneous_model", "layered_model", "device_count",
    "copy_bench", "SRC_REFERENCE_SIGMA", "SRC_FORCE", "SRC_EXPLOSIVE", "SRC_FORCE3",
    "WAVELET_GAUSSIAN", "WAVELET_DGAUSSIAN", "WAVELET_RICKER", "STORAGE_FP32", "STORAGE_FP16",
    "STORAGE_FP64", "COMPAT_XHALF_A", "COMPAT_ZHALF_MIN", "COMPAT_GHOST", "COMPAT_REFERENCE",
    "FIELD_NAMES", "Server", "SHOT_QUEUED", "SHOT_RUNNING", "SHOT_DONE", "anticline_model",
    "read_raw", "write_raw", "write_traces", "write_rechor", "run_noalloc", "NOALLOC_BYTES", "set_device_limit",
]

SRC_REFERENCE_SIGMA, SRC_FORCE, SRC_EXPLOSIVE, SRC_FORCE3 = 0, 1, 2, 3
WAVELET_GAUSSIAN, WAVELET_DGAUSSIAN, WAVELET_RICKER = 0, 1, 2
STORAGE_FP32, STORAGE_FP16, STORAGE_FP64 = 0, 1, 2
COMPAT_XHALF_A, COMPAT_ZHALF_MIN, COMPAT_GHOST, COMPAT_REFERENCE = 1, 2, 4, 7

FIELD_NAMES = [
    "vx", "vy", "vz", "sigmaxx", "sigmayy", "sigmazz", "sigmaxy", "sigmaxz", "sigmayz",
    "memory_dvx_dx", "memory_dvy_dy", "memory_dvz_dz", "memory_dvy_dx", "memory_dvx_dy",
    "memory_dvz_dx", "memory_dvx_dz", "memory_dvz_dy", "memory_dvy_dz",
    "memory_dsigmaxx_dx", "memory_dsigmaxy_dy", "memory_dsigmaxz_dz", "memory_dsigmaxy_dx",
    "memory_dsigmayy_dy", "memory_dsigmayz_dz", "memory_dsigmaxz_dx", "memory_dsigmayz_dy",
    "memory_dsigmazz_dz",
]


@dataclass
class Config:
    """The reference's parameter block (CPU:19-78, 146-153) with 0-based source/receivers."""
    nx: int = 101                # NIMX
    ny: int = 101                # NIMY
    nz: int = 101                # NIMZ
    dx: float = 10.0             # DELTAX
    dy: float = 10.0             # DELTAY
    dz: float = 10.0             # DELTAZ
    dt: float = 1.6e-3           # DELTATT
    nstep: int = 500             # NSTEP
    npoints_pml: int = 10        # NPOINTS_PML
    npower: float = 2.0          # NPOWER
    k_max_pml: float = 1.0       # K_MAX_PML
    alpha_max_pml: float = 0.0   # ALPHA_MAX_PML, <= 0 -> pi*f0
    rcoef: float = 1e-3          # Rcoef
    vpml: float = 3000.0         # vpml
    f0: float = 7.0
    t0: float = 0.0              # <= 0 -> 1.2/f0
    factor: float = 1e7
    angle_force: float = 0.0     # ANGLE_FORCE [deg]
    src_kind: int = SRC_FORCE
    wavelet: int = WAVELET_RICKER
    isrc: int = 50               # ISOURCE - 1
    jsrc: int = 50               # JSOURCE - 1
    ksrc: int = 50               # KSOURCE - 1
    receivers: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), dtype=np.int32))
    sample_every: int = 1        # sampgat
    energy_every: int = 0
    storage: int = STORAGE_FP32
    compat_flags: int = 0
    dirichlet: bool = True
    fp16_rescale_every: int = 100
    device: int = 0
    use_graph: bool = True
    block: int = 0
    zseg: int = 0
    free_surface: bool = False   # traction-free surface at the plane k = 0

    @property
    def t0_effective(self) -> float:
        return self.t0 if self.t0 > 0 else 1.2 / self.f0

    @property
    def nrec(self) -> int:
        return int(np.asarray(self.receivers).reshape(-1, 3).shape[0])

    @property
    def cells(self) -> int:
        return int(self.nx) * int(self.ny) * int(self.nz)

    def replace(self, **kw) -> "Config":
        return replace(self, **kw)

    @classmethod
    def from_reference(cls, NIMX, NIMY, NIMZ, DELTAX, DELTAT, NSTEP, NPOINTS_PML=10, f0=35.0,
                       factor=1e7, ISOURCE=None, JSOURCE=None, KSOURCE=11, ANGLE_FORCE=0.0,
                       sampgat=2, vpml=3000.0, **kw) -> "Config":
        """Build a Config from the reference's literal names (1-based source, CPU:20-72).

        Defaults reproduce the reference program: sigma source (CPU:620-622),
        derivative-of-Gaussian wavelet (CPU:610), all quirks (compat) on, no Dirichlet.
        """
        ISOURCE = (NIMX + 1) // 2 if ISOURCE is None else ISOURCE
        JSOURCE = (NIMY + 1) // 2 if JSOURCE is None else JSOURCE
        base = dict(nx=NIMX, ny=NIMY, nz=NIMZ, dx=DELTAX, dy=DELTAX, dz=DELTAX, dt=DELTAT,
                    nstep=NSTEP, npoints_pml=NPOINTS_PML, f0=f0, factor=factor,
                    isrc=ISOURCE - 1, jsrc=JSOURCE - 1, ksrc=KSOURCE - 1,
                    angle_force=ANGLE_FORCE, sample_every=sampgat, vpml=vpml,
                    src_kind=SRC_REFERENCE_SIGMA, wavelet=WAVELET_DGAUSSIAN,
                    compat_flags=COMPAT_REFERENCE, dirichlet=False)
        base.update(kw)
        return cls(**base)

    def to_struct(self):
        """C struct + keepalive for the receiver array."""
        rec = np.ascontiguousarray(np.asarray(self.receivers, dtype=np.int32).reshape(-1, 3))
        s = ConfigStruct()
        lib().ela3d_config_default(C.byref(s))
        for name in ("nx", "ny", "nz", "nstep", "npoints_pml", "src_kind", "wavelet", "isrc",
                     "jsrc", "ksrc", "sample_every", "energy_every", "storage",
                     "fp16_rescale_every", "device", "block", "zseg"):
            setattr(s, name, int(getattr(self, name)))
        for name in ("dx", "dy", "dz", "dt", "npower", "k_max_pml", "alpha_max_pml", "rcoef",
                     "vpml", "f0", "t0", "factor", "angle_force"):
            setattr(s, name, float(getattr(self, name)))
        s.compat_flags = int(self.compat_flags)
        s.dirichlet = 1 if self.dirichlet else 0
        s.use_graph = 1 if self.use_graph else 0
        s.free_surface = 1 if self.free_surface else 0
        s.nrec = rec.shape[0]
        s.rec = rec.ctypes.data_as(C.POINTER(C.c_int32)) if rec.size else C.POINTER(C.c_int32)()
        return s, rec

    def validate(self):
        s, keep = self.to_struct()
        check(lib().ela3d_validate(C.byref(s)))

    def plan(self) -> dict:
        """Host-side expansion of this config (no GPU): cells, C-PML regions, traffic."""
        s, keep = self.to_struct()
        inf = InfoStruct()
        check(lib().ela3d_plan(C.byref(s), C.byref(inf)))
        d = {k: getattr(inf, k) for k, _ in InfoStruct._fields_}
        d["region_lo"] = list(inf.region_lo)
        d["region_hi"] = list(inf.region_hi)
        return d

    def pml_tables(self, axis: int) -> dict:
        """The library's FP64 C-PML tables of one axis (d, K, alpha, a, b; full and half)."""
        s, keep = self.to_struct()
        n = (self.nx, self.ny, self.nz)[axis]
        out = np.zeros(10 * n, dtype=np.float64)
        check(lib().ela3d_pml_tables(C.byref(s), int(axis), _dptr(out), out.size))
        names = ["d", "K", "alpha", "a", "b", "d_half", "K_half", "alpha_half", "a_half", "b_half"]
        return {nm: out[i * n:(i + 1) * n].copy() for i, nm in enumerate(names)}


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Shot:
    """One forward shot resident on the GPU (ela3d_create ... ela3d_destroy)."""

    def __init__(self, cfg: Config, cp, cs, rho, slab=None):
        """slab=(k0, k1): a z-slab owning planes [k0, k1) of the grid cfg describes
        (ela3d_create_slab); step it with step_phase and halo_exchange."""
        self.cfg = cfg
        shape = (cfg.nz, cfg.ny, cfg.nx)
        self._model = [np.ascontiguousarray(np.asarray(a, dtype=np.float32)).reshape(shape)
                       for a in (cp, cs, rho)]
        s, keep = cfg.to_struct()
        h = C.c_void_p()
        self._nrec, self._nz, self.receiver_index = cfg.nrec, cfg.nz, None
        if slab is None:
            check(lib().ela3d_create(C.byref(s), _fptr(self._model[0]), _fptr(self._model[1]),
                                     _fptr(self._model[2]), C.byref(h)))
        else:
            k0, k1 = (int(v) for v in slab)
            idx = np.zeros(max(1, cfg.nrec), dtype=np.int32)
            nr = C.c_int32()
            check(lib().ela3d_create_slab(C.byref(s), k0, k1, _fptr(self._model[0]), _fptr(self._model[1]),
                                          _fptr(self._model[2]), C.byref(h),
                                          idx.ctypes.data_as(C.POINTER(C.c_int32)), idx.size, C.byref(nr)))
            self._nrec = nr.value
            self.receiver_index = idx[:nr.value].copy()  # global indices of the kept receivers
            self._nz = (k1 - k0) + (1 if k0 > 0 else 0) + (1 if k1 < cfg.nz else 0)
            self.koff = k0 - (1 if k0 > 0 else 0)
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().ela3d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, n: int = 1):
        check(lib().ela3d_step(self._h, int(n)))

    def sync(self):
        check(lib().ela3d_sync(self._h))

    def run(self):
        check(lib().ela3d_run(self._h))

    def info(self) -> dict:
        s = InfoStruct()
        check(lib().ela3d_info_get(self._h, C.byref(s)))
        d = {k: getattr(s, k) for k, _ in InfoStruct._fields_}
        d["region_lo"] = list(s.region_lo)
        d["region_hi"] = list(s.region_hi)
        return d

    def seismograms(self) -> np.ndarray:
        inf = self.info()
        out = np.zeros((inf["nsamples"], self._nrec, 3), dtype=np.float64)
        check(lib().ela3d_seismograms(self._h, _dptr(out), out.size))
        return out

    def energy(self) -> np.ndarray:
        inf = self.info()
        out = np.zeros((inf["nenergy"], 2), dtype=np.float64)
        check(lib().ela3d_energy(self._h, _dptr(out), out.size))
        return out

    def snapshot(self, fid: int) -> np.ndarray:
        out = np.zeros((self._nz, self.cfg.ny, self.cfg.nx), dtype=np.float64)
        check(lib().ela3d_snapshot(self._h, int(fid), _dptr(out), out.size))
        return out

    def set_field(self, fid: int, values):
        a = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
        check(lib().ela3d_set_field(self._h, int(fid), _dptr(a), a.size))

    def source_trace(self) -> np.ndarray:
        """[nstep+1][3]: sigma_xx, sigma_yy, sigma_zz at the source after each injection."""
        out = np.zeros((self.cfg.nstep + 1, 3), dtype=np.float64)
        check(lib().ela3d_source_trace(self._h, _dptr(out), out.size))
        return out

    def step_phase(self, phase: int):
        """Phase 0: the stress pass; phase 1: the velocity pass and the tail (one step)."""
        check(lib().ela3d_step_phase(self._h, int(phase)))

    def halo_bytes(self, phase: int, side: int) -> int:
        n = C.c_size_t()
        check(lib().ela3d_halo_bytes(self._h, int(phase), int(side), C.byref(n)))
        return n.value

    def halo_exchange(self, upper: "Shot", phase: int):
        """Both halos between this slab and the one above it after `phase` (same GPU)."""
        check(lib().ela3d_halo_exchange(self._h, upper._h, int(phase)))

    def checkpoint(self, path: str):
        """Write the whole device state and step count to `path`."""
        check(lib().ela3d_checkpoint_save(self._h, os.fsencode(path)))

    def restore(self, path: str):
        """Load a checkpoint written by a shot of the same config; continues bit-identically."""
        check(lib().ela3d_checkpoint_load(self._h, os.fsencode(path)))

    def profile(self, nsteps: int, every: int = 1):
        """Run nsteps; step s eagerly with hipEvents around every launch when
        s % every == 0, the others as graph replays.  Returns the per-launch
        totals of the instrumented steps and the event wall time of all nsteps."""
        st = (KernelStats * 8)()
        nk = C.c_int32()
        wall = C.c_double()
        check(lib().ela3d_profile_sampled(self._h, int(nsteps), int(every), st, 8, C.byref(nk),
                                          C.byref(wall)))
        ks = [dict(name=st[q].name.decode(), launches=st[q].launches, total_ms=st[q].total_ms,
                   alg_bytes_per_launch=st[q].alg_bytes_per_launch) for q in range(nk.value)]
        return ks, wall.value


def run_noalloc(cfg: Config, cp, cs, rho):
    """A whole shot in one call without device allocation (ela3d_run_noalloc): the
    device state lives in the library's static device buffer (NOALLOC_BYTES per GPU).
    Returns (seismograms [nsamples][nrec][3], energy [nenergy][2])."""
    shape = (cfg.nz, cfg.ny, cfg.nx)
    model = [np.ascontiguousarray(np.asarray(a, dtype=np.float32)).reshape(shape) for a in (cp, cs, rho)]
    p = cfg.plan()
    seis = np.zeros((p["nsamples"], cfg.nrec, 3), dtype=np.float64)
    en = np.zeros((p["nenergy"], 2), dtype=np.float64)
    s, keep = cfg.to_struct()
    check(lib().ela3d_run_noalloc(C.byref(s), _fptr(model[0]), _fptr(model[1]), _fptr(model[2]),
                                  _dptr(seis), seis.size, _dptr(en), en.size, None))
    return seis, en


NOALLOC_BYTES = 128 * 1024 * 1024  # ELA3D_NOALLOC_BYTES

SHOT_QUEUED, SHOT_RUNNING, SHOT_DONE = 0, 1, 2


class Server:
    """Continuous-batching shot server on one GPU (ela3d_server_create ... destroy).

    ``submit`` copies the configuration and model and returns a ticket; a device
    thread admits shots in arrival order within ``hbm_budget_bytes`` and
    ``max_active``, steps them concurrently (one stream per shot, graph replays in
    rounds of ``steps_per_round``) and retires them as they finish.
    ``plain_loop=True`` runs one shot at a time with per-step synchronous host
    traffic instead (the reference's way of driving its kernels).
    """

    def __init__(self, device: int = 0, hbm_budget_bytes: Optional[int] = None,
                 max_active: Optional[int] = None, steps_per_round: Optional[int] = None,
                 use_graph: bool = True, plain_loop: bool = False, batch_cells: Optional[int] = None):
        from ._lib import ServerOpts
        L = lib()
        o = ServerOpts()
        L.ela3d_server_opts_default(C.byref(o))
        o.device = device
        if hbm_budget_bytes is not None:
            o.hbm_budget_bytes = int(hbm_budget_bytes)
        if max_active is not None:
            o.max_active = int(max_active)
        if steps_per_round is not None:
            o.steps_per_round = int(steps_per_round)
        o.use_graph = 1 if use_graph else 0
        o.plain_loop = 1 if plain_loop else 0
        if batch_cells is not None:
            o.batch_cells = int(batch_cells)
        self.opts = o
        h = C.c_void_p()
        check(L.ela3d_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self._cfg = {}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().ela3d_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, cfg: Config, cp, cs, rho) -> int:
        shape = (cfg.nz, cfg.ny, cfg.nx)
        m = [np.ascontiguousarray(np.asarray(a, dtype=np.float32)).reshape(shape) for a in (cp, cs, rho)]
        s, keep = cfg.to_struct()
        t = C.c_uint64()
        check(lib().ela3d_submit(self._h, C.byref(s), _fptr(m[0]), _fptr(m[1]), _fptr(m[2]), C.byref(t)))
        self._cfg[t.value] = cfg
        return t.value

    def poll(self, ticket: int) -> int:
        st = C.c_int32()
        check(lib().ela3d_poll(self._h, C.c_uint64(ticket), C.byref(st)))
        return st.value

    def wait(self, ticket: int, timeout: float = -1.0) -> dict:
        """Block until the shot is done; returns its timing and status (0 = OK)."""
        from ._lib import ResultStruct
        r = ResultStruct()
        check(lib().ela3d_wait(self._h, C.c_uint64(ticket), float(timeout), C.byref(r)))
        return {k: getattr(r, k) for k, _ in ResultStruct._fields_}

    def result(self, ticket: int, release: bool = True):
        """(seismograms [nsamples][nrec][3], energy [nenergy][2], info) of a finished shot."""
        info = self.wait(ticket)
        seis = np.zeros((info["nsamples"], info["nrec"], 3), dtype=np.float64)
        en = np.zeros((info["nenergy"], 2), dtype=np.float64)
        check(lib().ela3d_result_seismograms(self._h, C.c_uint64(ticket), _dptr(seis), seis.size))
        check(lib().ela3d_result_energy(self._h, C.c_uint64(ticket), _dptr(en), en.size))
        if release:
            self.release(ticket)
        return seis, en, info

    def release(self, ticket: int):
        check(lib().ela3d_result_release(self._h, C.c_uint64(ticket)))
        self._cfg.pop(ticket, None)

    def stats(self) -> dict:
        out = (C.c_int64 * 10)()
        check(lib().ela3d_server_stats(self._h, out))
        keys = ["submitted", "done", "failed", "active", "queued", "bytes_in_use", "peak_bytes", "rounds",
                "batched", "batch_rounds"]
        return dict(zip(keys, list(out)))


def autotune(cfg: Config, model, steps: int = 8) -> dict:
    """Pick the fastest launch shape for this grid (cached for later shots)."""
    from ._lib import TuningStruct
    shape = (cfg.nz, cfg.ny, cfg.nx)
    m = [np.ascontiguousarray(np.asarray(a, dtype=np.float32)).reshape(shape) for a in model]
    s, keep = cfg.to_struct()
    t = TuningStruct()
    check(lib().ela3d_autotune(C.byref(s), _fptr(m[0]), _fptr(m[1]), _fptr(m[2]), int(steps),
                               C.byref(t)))
    return {k: getattr(t, k) for k, _ in TuningStruct._fields_}


def set_tuning(cfg: Config, block: int, zseg_interior: int, zseg_slab: int, vec: int = 1,
               block_slab: int = 0, vec_slab: int = 1, overlap: bool = False, merged: bool = False,
               dbuf: bool = False, fused: bool = False, tile: int = 1024, strips: int = 0):
    """Pin the launch shape later shots of this grid use (as if the tuner had chosen it).

    ``block``/``zseg_interior``/``vec`` shape the interior launches,
    ``block_slab`` (0 = ``block``)/``zseg_slab``/``vec_slab`` the absorbing-layer ones;
    ``merged`` runs one mixed launch per pass with the interior shape instead;
    ``overlap`` runs the absorbing-layer launches on a second stream beside the
    interior ones; ``dbuf`` keeps two field sets (each step reads one, writes the
    other; FP16 storage always updates in place and ignores it); ``fused`` (implies
    ``dbuf``) runs the fused interior step with ``tile``
    cells per workgroup and ``zseg_interior`` planes per segment, its workgroups
    walking ``strips`` xy strips of tiles (0: the default).  A shape a grid
    cannot run (a fused tile that does not fit, FP16 storage with ``fused``) falls
    back to the two-pass schedule when the shot is created."""
    from ._lib import TuningStruct
    s, keep = cfg.to_struct()
    t = TuningStruct(block=block, zseg_interior=zseg_interior, zseg_slab=zseg_slab, vec=vec,
                     candidates=0, ms_per_step=0.0, block_slab=block_slab, vec_slab=vec_slab,
                     overlap=1 if overlap else 0, merged=1 if merged else 0,
                     dbuf=1 if (dbuf or fused) else 0, fused=1 if fused else 0, tile=int(tile),
                     strips=int(strips))
    check(lib().ela3d_tuning_set(C.byref(s), C.byref(t)))


def set_device_limit(nbytes: int):
    """Cap the device bytes the library holds at once (failure injection: allocations
    past it fail with ELA3D_E_OOM); 0 removes the cap."""
    check(lib().ela3d_set_device_limit(int(nbytes)))


def device_count() -> int:
    n = C.c_int32()
    lib().ela3d_device_count(C.byref(n))
    return n.value


def copy_bench(nbytes: int, reps: int = 20, device: int = 0, variant: int = 0) -> float:
    """ms per device-to-device copy of nbytes with the library's stream-copy kernel."""
    ms = C.c_double()
    check(lib().ela3d_copy_bench(int(device), int(nbytes), int(reps), int(variant), C.byref(ms)))
    return ms.value


def homogeneous_model(nx, ny, nz, vp=3300.0, vs=3300.0 / 1.732, rho=2800.0):
    shape = (nz, ny, nx)
    cp = np.empty(shape, np.float32)
    cs = np.empty(shape, np.float32)
    rh = np.empty(shape, np.float32)
    check(lib().ela3d_model_homogeneous(nx, ny, nz, vp, vs, rho, _fptr(cp), _fptr(cs), _fptr(rh)))
    return cp, cs, rh


def layered_model(nx=401, ny=401, nz=301, h=10.0, seed=None, **kw):
    """Seeded 6-layer anticline model; returns (cp, cs, rho, vp_max)."""
    s = LayeredSpec()
    lib().ela3d_layered_default(C.byref(s))
    s.nx, s.ny, s.nz, s.h = nx, ny, nz, h
    if seed is not None:
        s.seed = int(seed)
    for k, v in kw.items():
        setattr(s, k, v)
    shape = (nz, ny, nx)
    cp = np.empty(shape, np.float32)
    cs = np.empty(shape, np.float32)
    rh = np.empty(shape, np.float32)
    vmax = C.c_double()
    check(lib().ela3d_model_layered(C.byref(s), _fptr(cp), _fptr(cs), _fptr(rh), C.byref(vmax)))
    return cp, cs, rh, vmax.value


def courant_dt(h: float, vp_max: float, courant: float = 0.5) -> float:
    """dt = courant * h / (vp_max * sqrt(3)) for the 3-D staggered scheme."""
    return courant * h / (vp_max * math.sqrt(3.0))


def anticline_model(n: int = 201):
    """The MATLAB anticline model (generate_3d_model_anticline.m) on an n^3 grid."""
    shape = (n, n, n)
    cp = np.empty(shape, np.float32)
    cs = np.empty(shape, np.float32)
    rh = np.empty(shape, np.float32)
    check(lib().ela3d_model_anticline(int(n), _fptr(cp), _fptr(cs), _fptr(rh)))
    return cp, cs, rh


def read_raw(path: str, shape, bytes_per_value: int = 8) -> np.ndarray:
    """Raw x-fastest model file (float64 or float32) as a float32 array of `shape` (nz, ny, nx)."""
    out = np.empty(tuple(shape), np.float32)
    check(lib().ela3d_read_raw(os.fsencode(path), out.size, int(bytes_per_value), _fptr(out)))
    return out


def write_raw(path: str, values, bytes_per_value: int = 4):
    a = np.ascontiguousarray(np.asarray(values, dtype=np.float32))
    check(lib().ela3d_write_raw(os.fsencode(path), a.size, int(bytes_per_value), _fptr(a)))


def write_traces(path: str, values, stride: int = 1, offset: int = 0, append: bool = False,
                 final_newline: bool = True):
    """Rows of values in the reference's text-trace layout (see ela3d_write_traces)."""
    a = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
    rows = a.shape[0] if a.ndim > 0 else 1
    cols = (a.size // max(rows, 1)) // stride
    check(lib().ela3d_write_traces(os.fsencode(path), _dptr(a), rows, cols, int(stride), int(offset),
                                   1 if append else 0, 1 if final_newline else 0))


def write_rechor(directory: str, seismograms, append: bool = True):
    """rechorvx.txt, rechorvy.txt, rechorvz.txt from seismograms [nsamples][nrec][3]."""
    for comp, name in enumerate(("rechorvx.txt", "rechorvy.txt", "rechorvz.txt")):
        write_traces(os.path.join(directory, name), seismograms, stride=3, offset=comp, append=append)
