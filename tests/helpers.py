"""Shared test helpers: seeded states, GPU <-> oracle transfer, error norms."""
from __future__ import annotations

import numpy as np

import oracle
from ela3d_amd import (COMPAT_REFERENCE, SRC_FORCE, WAVELET_RICKER, Config, homogeneous_model)


def small_cfg(nx=29, ny=23, nz=26, npml=4, compat=0, **kw) -> Config:
    """A small odd-sized grid with every code path present (x/y/z C-PML, interior)."""
    rec = np.asarray([(i, ny // 2, nz // 3) for i in range(0, nx, 3)] +
                     [(nx // 2, j, (2 * nz) // 3) for j in range(0, ny, 4)], dtype=np.int32)
    base = dict(nx=nx, ny=ny, nz=nz, dx=10.0, dy=12.0, dz=9.0, dt=1.0e-3, nstep=40,
                npoints_pml=npml, f0=18.0, factor=1e7, src_kind=SRC_FORCE,
                wavelet=WAVELET_RICKER, isrc=nx // 2, jsrc=ny // 2, ksrc=nz // 2, receivers=rec,
                energy_every=1, compat_flags=compat, dirichlet=(compat != COMPAT_REFERENCE))
    base.update(kw)
    return Config(**base)


def hetero_model(cfg: Config, seed=1):
    """Smoothly varying seeded model so material averages matter."""
    rng = np.random.default_rng(seed)
    shape = (cfg.nz, cfg.ny, cfg.nx)
    cp = (3000.0 + 400.0 * rng.random(shape)).astype(np.float32)
    cs = (cp / (1.6 + 0.2 * rng.random(shape))).astype(np.float32)
    rho = (2200.0 + 600.0 * rng.random(shape)).astype(np.float32)
    return cp, cs, rho


def homo(cfg: Config):
    return homogeneous_model(cfg.nx, cfg.ny, cfg.nz, 3300.0, 3300.0 / 1.732, 2800.0)


def random_state(cfg: Config, seed=0, dtype=np.float32) -> oracle.State:
    """Random wavefields; random memory variables wherever their own table has a != 0."""
    rng = np.random.default_rng(seed)
    st = oracle.State(cfg, dtype)
    shape = (cfg.nz, cfg.ny, cfg.nx)
    for q in range(9):
        scale = 1.0 if q < 3 else 1.0e6
        st.f[q][...] = (scale * rng.standard_normal(shape)).astype(dtype)
    tabs = [oracle.tables(cfg, a) for a in range(3)]
    for q in range(18):
        axis, half = oracle.MEM_TABLE[q]
        a = tabs[axis]["a_half" if half else "a"]
        mask1 = a != 0.0
        if axis == 0:
            mask = np.broadcast_to(mask1[None, None, :], shape)
        elif axis == 1:
            mask = np.broadcast_to(mask1[None, :, None], shape)
        else:
            mask = np.broadcast_to(mask1[:, None, None], shape)
        scale = 1.0e-3 if q < 9 else 1.0e3
        vals = (scale * rng.standard_normal(shape)).astype(dtype)
        st.m[q][...] = np.where(mask, vals, 0).astype(dtype)
    return st


def push_state(shot, st: oracle.State):
    for q in range(27):
        shot.set_field(q, st.field(q).astype(np.float64))


def max_rel(a, b):
    d = np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)))
    s = np.max(np.abs(np.asarray(b, np.float64)))
    return d / s if s > 0 else d


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / nb if nb > 0 else np.linalg.norm(a.ravel())


def compare_state(shot, st: oracle.State, tol, fields=range(27)):
    """Largest per-field max-abs error relative to the field's max; returns (worst, name)."""
    worst, wname = 0.0, ""
    for q in fields:
        g = shot.snapshot(q)
        r = st.field(q).astype(np.float64)
        e = max_rel(g, r)
        if e > worst:
            worst, wname = e, (oracle.FIELD_NAMES + oracle.MEM_NAMES)[q]
        assert e <= tol, f"field {q} ({(oracle.FIELD_NAMES + oracle.MEM_NAMES)[q]}): {e:.3e} > {tol}"
    return worst, wname
