"""GPU: the largest grid a shot accepts, against a small grid, by translation invariance.

A padded field array must stay below 4 GiB, so that every access is the array's
base plus a 32-bit byte offset: 2^30 elements in FP32 and FP16 storage, 2^29 in
FP64.  1024 x 1024 x 1023 (1,072,695,424 padded elements) and, for FP64,
1024 x 1024 x 511 are just inside those limits.  With the source placed 24 cells
from the far corner, the highest offsets of every array take part in the first
steps.

In a homogeneous medium, and before the wave reaches any absorbing layer or face,
the seismograms of receivers around the source depend only on their offsets from
it.  The same source and receiver offsets on a 48^3 grid must therefore give the
same samples bit for bit, as long as both grids run the same kernel instantiation
(same launch shape) over cells of the interior box.  In FP16 storage the per-field
scales follow the fields' maxima, which sit around the source on both grids, so
they match too.
"""
import numpy as np
import pytest

from ela3d_amd import (Config, SRC_FORCE, STORAGE_FP16, STORAGE_FP32, STORAGE_FP64, WAVELET_RICKER, Shot,
                       homogeneous_model, set_tuning)

pytestmark = pytest.mark.gpu

OFFSETS = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, -1), (2, 1, -1), (-2, -2, 2)],
                   dtype=np.int32)
NSTEP = 6


def shot_cfg(nx, ny, nz, src, storage=STORAGE_FP32):
    rec = (np.asarray(src, dtype=np.int32)[None, :] + OFFSETS).astype(np.int32)
    return Config(nx=nx, ny=ny, nz=nz, dx=10.0, dy=10.0, dz=10.0, dt=1.6e-3, nstep=NSTEP, npoints_pml=10,
                  f0=7.0, factor=1e7, src_kind=SRC_FORCE, wavelet=WAVELET_RICKER, isrc=int(src[0]),
                  jsrc=int(src[1]), ksrc=int(src[2]), receivers=rec, energy_every=0,
                  storage=storage)


def run(cfg):
    set_tuning(cfg, block=128, zseg_interior=2, vec=1, block_slab=128, zseg_slab=2, vec_slab=1)
    model = homogeneous_model(cfg.nx, cfg.ny, cfg.nz)
    print(f"grid {cfg.nx}x{cfg.ny}x{cfg.nz}: model ready", flush=True)
    with Shot(cfg, *model) as shot:
        del model
        assert shot.info()["block"] == 128
        shot.run()
        return shot.seismograms()


@pytest.mark.parametrize("storage,nz", [(STORAGE_FP32, 1023), (STORAGE_FP16, 1023), (STORAGE_FP64, 511)])
def test_largest_grid_matches_small_grid_bit_for_bit(gpu, storage, nz):
    small = shot_cfg(48, 48, 48, (24, 24, 24), storage)
    big = shot_cfg(1024, 1024, nz, (1024 - 24, 1024 - 24, nz - 24), storage)
    plan = big.plan()
    print(f"largest grid: {plan['device_bytes'] / 1e9:.1f} GB on the device", flush=True)
    s_small = run(small)
    s_big = run(big)
    assert np.abs(s_small).max() > 0
    np.testing.assert_array_equal(s_big, s_small)
