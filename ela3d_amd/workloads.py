"""The named workloads: configs, receiver layouts and models with their sizes.

* ``homogeneous_cube``: 101^3 homogeneous cube, dx=10 m, cp=3300, cs=cp/1.732,
  rho=2800, dt=1.6e-3 s, NSTEP=500, 10-point C-PML, Ricker force in vy at
  f0=7 Hz at the centre, two receiver lines of 21 stations along x at 1/4 and
  3/4 depth (i0 = 0,5,...,100 as the reference's ``i += xlen`` walk, CPU:642).
* ``classic``: the SEISMIC_CPML shape 101 x 641 x 640, same medium and source,
  NSTEP=2500, receivers every 10 points along y at the first plane below the
  top C-PML (k0 = NPOINTS_PML), i0 = 50.
* ``layered``: seeded 6-layer anticline 401 x 401 x 301 at 10 m, dt at
  Courant 0.5 of the realised max Vp, NSTEP=3000, explosive source at f0=10 Hz
  20 m below the inner edge of the top C-PML, vpml = max Vp.
* ``serving_stream``: 256 forward shots arriving Poisson at 4 shots/s, each a
  seeded homogeneous (p = 0.5) or layered cube with edge uniform in
  {64, 96, 128, 192, 256} at 10 m, NSTEP uniform in 500..2000, Ricker force at
  f0 uniform in [5, 15] Hz at a random interior point, 64 random receivers,
  dt at Courant 0.5, FP32.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import (SRC_EXPLOSIVE, SRC_FORCE, STORAGE_FP16, STORAGE_FP32, WAVELET_RICKER, Config,
               courant_dt, homogeneous_model, layered_model)

CP, CS, RHO = 3300.0, 3300.0 / 1.732, 2800.0


def receiver_lines_x(nx, j, ks, step):
    """Stations along x at i = 0, step, ... < nx for each depth in ks."""
    r = [(i, j, k) for k in ks for i in range(0, nx, step)]
    return np.asarray(r, dtype=np.int32)


def homogeneous_cube(nstep=500, n=101, **kw):
    c = n // 2
    rec = receiver_lines_x(n, c, [n // 4, (3 * n) // 4], max(1, (n - 1) // 20))
    cfg = Config(nx=n, ny=n, nz=n, dx=10.0, dy=10.0, dz=10.0, dt=1.6e-3, nstep=nstep,
                 npoints_pml=10, f0=7.0, factor=1e7, src_kind=SRC_FORCE, wavelet=WAVELET_RICKER,
                 isrc=c, jsrc=c, ksrc=c, receivers=rec, energy_every=1)
    cfg = cfg.replace(**kw)
    return cfg, homogeneous_model(cfg.nx, cfg.ny, cfg.nz, CP, CS, RHO)


def classic(nstep=2500, nx=101, ny=641, nz=640, **kw):
    rec = np.asarray([(nx // 2, j, 10) for j in range(0, ny, 10)], dtype=np.int32)
    cfg = Config(nx=nx, ny=ny, nz=nz, dx=10.0, dy=10.0, dz=10.0, dt=1.6e-3, nstep=nstep,
                 npoints_pml=10, f0=7.0, factor=1e7, src_kind=SRC_FORCE, wavelet=WAVELET_RICKER,
                 isrc=nx // 2, jsrc=ny // 2, ksrc=nz // 2, receivers=rec, energy_every=10,
                 storage=STORAGE_FP32)
    cfg = cfg.replace(**kw)
    return cfg, homogeneous_model(cfg.nx, cfg.ny, cfg.nz, CP, CS, RHO)


def layered(nstep=3000, nx=401, ny=401, nz=301, seed=20240611, fp16=False, **kw):
    cp, cs, rho, vmax = layered_model(nx, ny, nz, 10.0, seed=seed)
    dt = courant_dt(10.0, vmax, 0.5)
    ksrc = 12
    rec = np.asarray([(i, ny // 2, 10) for i in range(0, nx, 10)] +
                     [(nx // 2, j, 10) for j in range(0, ny, 10)], dtype=np.int32)
    cfg = Config(nx=nx, ny=ny, nz=nz, dx=10.0, dy=10.0, dz=10.0, dt=dt, nstep=nstep,
                 npoints_pml=10, f0=10.0, factor=1e7, vpml=vmax, src_kind=SRC_EXPLOSIVE,
                 wavelet=WAVELET_RICKER, isrc=nx // 2, jsrc=ny // 2, ksrc=ksrc, receivers=rec,
                 energy_every=10, storage=STORAGE_FP16 if fp16 else STORAGE_FP32)
    cfg = cfg.replace(**kw)
    return cfg, (cp, cs, rho)


SERVING_EDGES = (64, 96, 128, 192, 256)


@dataclass
class ShotSpec:
    """One request of the serving stream (the model is generated on demand)."""
    index: int
    arrival: float      # seconds after the stream starts
    edge: int
    nstep: int
    f0: float
    layered: bool
    model_seed: int
    src: tuple
    receivers: np.ndarray

    def materialize(self):
        """(Config, (cp, cs, rho)) of this shot."""
        n = self.edge
        if self.layered:
            depth = n * 10.0
            cp, cs, rho, vmax = layered_model(n, n, n, 10.0, seed=self.model_seed,
                                              bump_amp=0.15 * depth, bump_width=0.3 * depth)
            model = (cp, cs, rho)
        else:
            vmax = CP
            model = homogeneous_model(n, n, n, CP, CS, RHO)
        cfg = Config(nx=n, ny=n, nz=n, dx=10.0, dy=10.0, dz=10.0, dt=courant_dt(10.0, vmax, 0.5),
                     nstep=self.nstep, npoints_pml=10, f0=self.f0, factor=1e7, vpml=vmax,
                     src_kind=SRC_FORCE, wavelet=WAVELET_RICKER, isrc=self.src[0],
                     jsrc=self.src[1], ksrc=self.src[2], receivers=self.receivers,
                     energy_every=0, storage=STORAGE_FP32)
        return cfg, model

    @property
    def cell_updates(self) -> int:
        return self.edge ** 3 * self.nstep


def serving_stream(nshots=256, rate=4.0, seed=7, edges=SERVING_EDGES, nstep_range=(500, 2000),
                   f0_range=(5.0, 15.0), nrec=64, npml=10):
    """The seeded request stream: a list of ShotSpec in arrival order."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for q in range(nshots):
        t += rng.exponential(1.0 / rate)
        n = int(rng.choice(edges))
        lo, hi = npml + 2, n - npml - 2  # strictly inside the absorbing layers
        src = tuple(int(v) for v in rng.integers(lo, hi, size=3))
        rec = rng.integers(lo, hi, size=(nrec, 3)).astype(np.int32)
        out.append(ShotSpec(index=q, arrival=t, edge=n,
                            nstep=int(rng.integers(nstep_range[0], nstep_range[1] + 1)),
                            f0=float(rng.uniform(*f0_range)), layered=bool(rng.random() < 0.5),
                            model_seed=int(rng.integers(1, 2**62)), src=src, receivers=rec))
    return out


def fifo_latencies(arrivals, service):
    """Latency of each request on a one-at-a-time FIFO server (Lindley's recursion):
    start = max(arrival, previous completion), latency = start + service - arrival."""
    done = 0.0
    out = []
    for a, s in zip(arrivals, service):
        start = max(a, done)
        done = start + s
        out.append(done - a)
    return np.asarray(out)
