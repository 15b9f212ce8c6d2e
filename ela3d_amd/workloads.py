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
"""
from __future__ import annotations

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
