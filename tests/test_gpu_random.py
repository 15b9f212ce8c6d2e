"""Seeded randomized parity: grids, C-PML widths, quirk flags, source kinds, wavelets,
rigid faces, energy cadence, storage, receivers, launch shapes and schedules drawn at
random (fixed seeds), each compared with the oracle after one step from a random
state; and FP16-storage shots from rest over a whole run against the FP64 oracle.
Every case's device address arithmetic is first checked by the CPU model of
tests/test_index_bounds.py; the CPU test below runs that check for all cases, and the
GPU test repeats it before launching anything."""
import numpy as np
import pytest

import oracle
from ela3d_amd import (COMPAT_GHOST, COMPAT_REFERENCE, COMPAT_XHALF_A, COMPAT_ZHALF_MIN, Ela3dError,
                       SRC_EXPLOSIVE, SRC_FORCE, SRC_FORCE3, SRC_REFERENCE_SIGMA, STORAGE_FP16,
                       STORAGE_FP32, STORAGE_FP64, WAVELET_DGAUSSIAN, WAVELET_GAUSSIAN, WAVELET_RICKER, Shot,
                       set_tuning)
from helpers import compare_state, hetero_model, max_rel, push_state, random_state, small_cfg
from test_index_bounds import fused_violations, violations

NCASES = 48
TOL = {STORAGE_FP32: 2e-6, STORAGE_FP64: 1e-12}
DT = {STORAGE_FP32: np.float32, STORAGE_FP64: np.float64}


def draw(seed):
    """One random case: (config, launch shape), regenerated until the config is valid."""
    rng = np.random.default_rng(1000 + seed)
    pick = lambda xs: xs[int(rng.integers(len(xs)))]
    while True:
        nx, ny, nz = int(rng.integers(9, 61)), int(rng.integers(9, 49)), int(rng.integers(9, 41))
        npml = int(rng.integers(0, 6))
        rec = np.stack([rng.integers(0, nx, 8), rng.integers(0, ny, 8), rng.integers(0, nz, 8)], 1)
        rec = rec[: int(rng.integers(0, 9))].astype(np.int32)
        cfg = small_cfg(nx=nx, ny=ny, nz=nz, npml=npml, nstep=4,
                        compat=pick([0, COMPAT_REFERENCE, COMPAT_XHALF_A, COMPAT_ZHALF_MIN, COMPAT_GHOST]),
                        src_kind=pick([SRC_REFERENCE_SIGMA, SRC_FORCE, SRC_EXPLOSIVE, SRC_FORCE3]),
                        wavelet=pick([WAVELET_GAUSSIAN, WAVELET_DGAUSSIAN, WAVELET_RICKER]),
                        storage=STORAGE_FP64 if rng.random() < 0.25 else STORAGE_FP32,
                        isrc=int(rng.integers(1, nx - 1)), jsrc=int(rng.integers(1, ny - 1)),
                        ksrc=int(rng.integers(1, nz - 1)), receivers=rec, angle_force=float(rng.uniform(0, 90)),
                        energy_every=pick([0, 1, 3]))
        cfg = cfg.replace(dirichlet=bool(rng.random() < 0.5))
        # a free surface in about a third of the cases (its own generator keeps the other draws)
        cfg = cfg.replace(free_surface=bool(np.random.default_rng(7000 + seed).random() < 0.35))
        try:
            cfg.validate()
        except Ela3dError:
            continue
        shape = dict(block=pick([64, 128, 256]), zseg_interior=pick([1, 2, 3, 4, 8]), vec=pick([1, 2, 4]),
                     block_slab=pick([64, 128, 256]), zseg_slab=pick([1, 2, 4]), vec_slab=pick([1, 2]),
                     overlap=bool(rng.random() < 0.3), merged=bool(rng.random() < 0.25),
                     dbuf=bool(rng.random() < 0.3))
        if rng.random() < 0.2 and cfg.storage in (STORAGE_FP32, STORAGE_FP64) and npml > 0:
            # FP64 tiles hold at most 512 cells; the tile order's strip count from a
            # generator of its own (0: the default order)
            tiles = [256, 512, 1024] if cfg.storage == STORAGE_FP32 else [256, 512, 512]
            strips = int(np.random.default_rng(9000 + seed).integers(0, 5))
            shape.update(fused=True, tile=pick(tiles), dbuf=True, merged=False, strips=strips)
        return cfg, shape


def address_violations(cfg, sh):
    """The CPU address model of the launches this case makes."""
    if sh.get("merged"):  # the absorbing-layer boxes run with the interior shape
        bad = violations(cfg, sh["block"], sh["vec"], sh["zseg_interior"], sh["zseg_slab"],
                         block_slab=sh["block"], vec_slab=sh["vec"])
    else:
        bad = violations(cfg, sh["block"], sh["vec"], sh["zseg_interior"], sh["zseg_slab"],
                         block_slab=sh["block_slab"], vec_slab=sh["vec_slab"])
    if sh.get("fused"):
        f = fused_violations(cfg, sh["tile"], sh["zseg_interior"])
        bad = bad + (f or [])
    return bad


def test_random_cases_address_model():
    for seed in range(NCASES):
        cfg, sh = draw(seed)
        assert address_violations(cfg, sh) == [], (seed, sh)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(NCASES))
def test_random_case_one_step(gpu, seed):
    cfg, sh = draw(seed)
    assert address_violations(cfg, sh) == []
    print(f"case {seed}: {cfg.nx}x{cfg.ny}x{cfg.nz} npml {cfg.npoints_pml} compat {cfg.compat_flags} "
          f"src {cfg.src_kind} wavelet {cfg.wavelet} storage {cfg.storage} nrec {cfg.nrec} "
          f"free surface {cfg.free_surface} {sh}", flush=True)
    set_tuning(cfg, **sh)
    model = hetero_model(cfg, seed=seed)
    st = random_state(cfg, seed=seed, dtype=DT[cfg.storage])
    with Shot(cfg, *model) as shot:
        fused = shot.info()["fused"]
        push_state(shot, st)
        shot.step(1)
        shot.sync()
        seis_o, en_o, st = oracle.run(cfg, model, 1, st=st, dtype=DT[cfg.storage])
        compare_state(shot, st, TOL[cfg.storage])
        seis_g, en_g = shot.seismograms()[:1], shot.energy()[:1]
    if cfg.nrec:
        assert max_rel(seis_g, seis_o[:1]) <= TOL[cfg.storage]
    assert np.allclose(en_g, en_o[:1], rtol=10 * TOL[cfg.storage], atol=0)
    assert fused == int(sh.get("fused", False)) or fused == 0  # refused fused shapes fall back to two passes


NFP16 = 8


def draw_fp16(seed, storage=STORAGE_FP16):
    """One random multi-step case from rest: (config, launch shape), FP16 storage
    unless `storage` says otherwise.  Every source with every wavelet: a Gaussian
    driving a stress source leaves a static stress at the source cell 10^4-10^5
    times the wave's, which FP16 storage keeps in an FP32 side slot."""
    rng = np.random.default_rng((5000 if storage == STORAGE_FP16 else 6000) + seed)
    pick = lambda xs: xs[int(rng.integers(len(xs)))]
    while True:
        nx, ny, nz = (int(rng.integers(24, 57)) for _ in range(3))
        npml = int(rng.integers(2, 9))
        lo = npml + 1
        rec = np.stack([rng.integers(1, nx - 1, 8), rng.integers(1, ny - 1, 8), rng.integers(1, nz - 1, 8)], 1)
        src = pick([SRC_REFERENCE_SIGMA, SRC_FORCE, SRC_EXPLOSIVE, SRC_FORCE3])
        cfg = small_cfg(nx=nx, ny=ny, nz=nz, npml=npml, nstep=int(rng.integers(150, 251)),
                        compat=pick([0, COMPAT_REFERENCE]), storage=storage, src_kind=src,
                        wavelet=pick([WAVELET_GAUSSIAN, WAVELET_DGAUSSIAN, WAVELET_RICKER]),
                        isrc=int(rng.integers(lo, nx - lo)), jsrc=int(rng.integers(lo, ny - lo)),
                        ksrc=int(rng.integers(lo, nz - lo)), receivers=rec.astype(np.int32),
                        angle_force=float(rng.uniform(0, 90)), energy_every=pick([1, 5]))
        cfg = cfg.replace(free_surface=bool(np.random.default_rng(8000 + seed).random() < 0.35))
        try:
            cfg.validate()
        except Ela3dError:
            continue
        shape = dict(block=pick([64, 128, 256]), zseg_interior=pick([1, 2, 4, 8]), vec=pick([1, 2]),
                     block_slab=pick([64, 128, 256]), zseg_slab=pick([1, 2, 4]), vec_slab=pick([1, 2]),
                     overlap=bool(rng.random() < 0.3))
        return cfg, shape


def test_random_fp16_cases_address_model():
    for seed in range(NFP16):
        for storage in (STORAGE_FP16, STORAGE_FP32):
            cfg, sh = draw_fp16(seed, storage)
            assert address_violations(cfg, sh) == [], (seed, storage, sh)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(NFP16))
def test_random_fp16_case_many_steps(gpu, seed):
    """FP16 storage over a whole shot from rest, against the FP64 oracle: finite
    everywhere, seismograms and total energy within the FP16 bar (3e-2)."""
    cfg, sh = draw_fp16(seed)
    assert address_violations(cfg, sh) == []
    print(f"fp16 case {seed}: {cfg.nx}x{cfg.ny}x{cfg.nz} npml {cfg.npoints_pml} compat {cfg.compat_flags} "
          f"src {cfg.src_kind} wavelet {cfg.wavelet} nstep {cfg.nstep} {sh}", flush=True)
    set_tuning(cfg, **sh)
    model = hetero_model(cfg, seed=100 + seed)
    with Shot(cfg, *model) as shot:
        shot.run()
        s16, e16 = shot.seismograms(), shot.energy()
    s64, e64, _ = oracle.run(cfg, model, dtype=np.float64)
    err = float(np.linalg.norm(s16 - s64) / np.linalg.norm(s64))
    eerr = float(np.max(np.abs(e16.sum(1) - e64.sum(1))) / np.max(e64.sum(1)))
    print(f"fp16 case {seed}: seismogram rel L2 {err:.3e}, max energy deviation {eerr:.3e}")
    assert np.all(np.isfinite(s16)) and np.all(np.isfinite(e16))
    assert err < 3e-2 and eerr < 3e-2


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(NFP16))
def test_random_fp32_case_many_steps(gpu, seed):
    """The same kind of whole shot from rest in FP32 storage: seismograms within the
    1e-4 relative L2 bar of the FP64 oracle (every wavelet with every source)."""
    cfg, sh = draw_fp16(seed, STORAGE_FP32)
    assert address_violations(cfg, sh) == []
    print(f"fp32 case {seed}: {cfg.nx}x{cfg.ny}x{cfg.nz} npml {cfg.npoints_pml} compat {cfg.compat_flags} "
          f"src {cfg.src_kind} wavelet {cfg.wavelet} nstep {cfg.nstep} {sh}", flush=True)
    set_tuning(cfg, **sh)
    model = hetero_model(cfg, seed=200 + seed)
    with Shot(cfg, *model) as shot:
        shot.run()
        s32, e32 = shot.seismograms(), shot.energy()
    s64, e64, _ = oracle.run(cfg, model, dtype=np.float64)
    err = float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64))
    eerr = float(np.max(np.abs(e32.sum(1) - e64.sum(1))) / np.max(e64.sum(1)))
    print(f"fp32 case {seed}: seismogram rel L2 {err:.3e}, max energy deviation {eerr:.3e}")
    assert err <= 1e-4 and eerr <= 1e-4


def test_fp16_explosive_gaussian_is_refused():
    """The explosive source with the Gaussian leaves static stresses around the source
    in all six stresses; FP16 storage refuses the combination (every other source and
    wavelet is drawn by the randomized FP16 cases)."""
    cfg = small_cfg(nx=24, ny=26, nz=46, npml=4, storage=STORAGE_FP16, src_kind=SRC_EXPLOSIVE,
                    wavelet=WAVELET_GAUSSIAN)
    with pytest.raises(Ela3dError) as e:
        cfg.validate()
    assert e.value.code == -1 and "static stress" in str(e.value)
    cfg.replace(storage=STORAGE_FP32).validate()


@pytest.mark.gpu
@pytest.mark.parametrize("src", [SRC_REFERENCE_SIGMA])
def test_fp16_stress_source_with_gaussian(gpu, src):
    """The case FP16 storage once failed: a Gaussian (non-zero mean) driving the
    reference's sigma source on the 24x26x46 grid (seismogram rel L2 4.7e-2 before
    the source cell's normal stresses moved to the FP32 side slot; 1.45e-2 after).
    Against the FP64 oracle within the FP16 bar, and the source cell's stresses in
    the snapshot are the side slot's, close to the oracle's."""
    cfg = small_cfg(nx=24, ny=26, nz=46, npml=4, nstep=220, storage=STORAGE_FP16, src_kind=src,
                    wavelet=WAVELET_GAUSSIAN, isrc=11, jsrc=13, ksrc=20, energy_every=5,
                    receivers=np.asarray([(i, 13, k) for i in range(2, 22, 3) for k in (10, 30)], np.int32))
    set_tuning(cfg, block=128, zseg_interior=2, vec=1, block_slab=128, zseg_slab=2, vec_slab=1)
    model = hetero_model(cfg, seed=77)
    with Shot(cfg, *model) as shot:
        shot.run()
        s16, e16 = shot.seismograms(), shot.energy()
        snap = [shot.snapshot(q) for q in (3, 4, 5)]
    s64, e64, st = oracle.run(cfg, model, dtype=np.float64)
    err = float(np.linalg.norm(s16 - s64) / np.linalg.norm(s64))
    eerr = float(np.max(np.abs(e16.sum(1) - e64.sum(1))) / np.max(e64.sum(1)))
    print(f"fp16 stress source {src} with a Gaussian: seismogram rel L2 {err:.3e}, energy {eerr:.3e}")
    assert err < 3e-2 and eerr < 3e-2
    k, j, i = cfg.ksrc, cfg.jsrc, cfg.isrc
    big = max(abs(st.f[q][k, j, i]) for q in (3, 4, 5))  # the injected static stress
    for q, sn in zip((3, 4, 5), snap):
        assert abs(sn[k, j, i] - st.f[q][k, j, i]) <= 1e-3 * big, q

