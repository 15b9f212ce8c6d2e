"""The allocation-free entry point (ela3d_run_noalloc) against the allocating one.

ela3d_run_noalloc runs a whole shot with its device state in the library's static
device buffer instead of memory from hipMalloc.  With the same launch shape it must
give exactly what ela3d_create + ela3d_run give: the same kernels, the same order of
every floating-point operation.  Two calls in a row, the second on a smaller grid,
check that the buffer's previous contents never leak into a run; a grid too large
for the buffer is refused before anything runs.
"""
import numpy as np
import pytest

import ela3d_amd
from ela3d_amd import (SRC_EXPLOSIVE, SRC_FORCE, STORAGE_FP32, WAVELET_RICKER, Config, Ela3dError,
                       Shot)

pytestmark = pytest.mark.gpu


def _model(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    shape = (nz, ny, nx)
    cp = (3000 + 600 * rng.random(shape)).astype(np.float32)
    cs = (cp / (1.6 + 0.2 * rng.random(shape))).astype(np.float32)
    rho = (2100 + 500 * rng.random(shape)).astype(np.float32)
    return cp, cs, rho


def _cfg(nx, ny, nz, src, kind, nstep):
    rec = np.asarray([(i, ny // 2, nz // 3) for i in range(1, nx - 1, 3)], dtype=np.int32)
    return Config(nx=nx, ny=ny, nz=nz, dt=1.2e-3, nstep=nstep, npoints_pml=5, f0=14.0,
                  src_kind=kind, wavelet=WAVELET_RICKER, isrc=src[0], jsrc=src[1], ksrc=src[2],
                  receivers=rec, energy_every=1, storage=STORAGE_FP32, block=128, zseg=2)


def _allocating(cfg, model):
    with Shot(cfg, *model) as s:
        s.run()
        return s.seismograms(), s.energy()


def test_noalloc_equals_create_run_twice_second_smaller(gpu):
    cases = [
        (_cfg(52, 44, 40, (25, 21, 19), SRC_FORCE, 120), 7),
        (_cfg(30, 27, 23, (14, 12, 11), SRC_EXPLOSIVE, 90), 8),  # smaller, after the larger
    ]
    for cfg, seed in cases:
        model = _model(cfg.nx, cfg.ny, cfg.nz, seed)
        print(f"noalloc case {cfg.nx}x{cfg.ny}x{cfg.nz} seed {seed}", flush=True)
        seis, en = ela3d_amd.run_noalloc(cfg, *model)
        ref_seis, ref_en = _allocating(cfg, model)
        assert np.any(ref_seis) and np.all(np.isfinite(seis))
        assert np.array_equal(seis, ref_seis)
        assert np.array_equal(en, ref_en)


def test_noalloc_refuses_a_grid_beyond_the_buffer(gpu):
    n = 200  # 27 FP32 arrays of 8e6 cells need far more than the buffer
    cfg = _cfg(n, n, n, (n // 2, n // 2, n // 2), SRC_FORCE, 2)
    one = np.ones((n, n, n), np.float32)
    with pytest.raises(Ela3dError) as e:
        ela3d_amd.run_noalloc(cfg, one * 3000, one * 1800, one * 2200)
    assert e.value.code == -6  # ELA3D_E_RANGE
    assert "available" in str(e.value)
