"""GPU: FP16 storage with per-field power-of-two scales, the layered anticline
model, and FP32 against the GPU FP64 instantiation at a larger size.

FP16 storage keeps an 11-bit significand, so its seismogram error against the
FP64 oracle is reported (printed) and bounded loosely at 3e-2 relative L2; the
FP32 runs keep the 1e-4 bar of the workload.
"""
import numpy as np
import pytest

import oracle
from ela3d_amd import (STORAGE_FP16, STORAGE_FP32, STORAGE_FP64, Shot, layered_model, workloads)
from helpers import hetero_model, rel_l2, small_cfg

pytestmark = pytest.mark.gpu


def run(cfg, model):
    with Shot(cfg, *model) as shot:
        shot.run()
        return shot.seismograms(), shot.energy()


def test_fp16_storage_small_grid(gpu):
    cfg = small_cfg(nstep=200, storage=STORAGE_FP16)
    model = hetero_model(cfg, seed=12)
    s16, e16 = run(cfg, model)
    s64, e64, _ = oracle.run(cfg, model, dtype=np.float64)
    err = rel_l2(s16, s64)
    eerr = np.max(np.abs(e16.sum(1) - e64.sum(1))) / np.max(e64.sum(1))
    print(f"FP16 storage, small grid: seismogram rel L2 {err:.3e}, max energy deviation {eerr:.3e}")
    assert np.all(np.isfinite(s16))
    assert err < 3e-2 and eerr < 3e-2


def test_fp16_storage_homogeneous_cube(gpu):
    cfg, model = workloads.homogeneous_cube(storage=STORAGE_FP16)
    s16, e16 = run(cfg, model)
    s64, e64, _ = oracle.run(cfg, model, dtype=np.float64, threads=16)
    err = rel_l2(s16, s64)
    eerr = np.max(np.abs(e16.sum(1) - e64.sum(1))) / np.max(e64.sum(1))
    print(f"FP16 storage, 101^3: seismogram rel L2 {err:.3e}, max energy deviation {eerr:.3e}")
    assert err < 3e-2 and eerr < 3e-2


def test_fp16_snapshot_round_trip(gpu):
    """set_field/snapshot honour the live per-field scale."""
    cfg = small_cfg(nstep=30, storage=STORAGE_FP16)
    model = hetero_model(cfg, seed=3)
    with Shot(cfg, *model) as shot:
        shot.step(25)
        shot.sync()
        a = shot.snapshot(3)
        shot.set_field(3, a)
        b = shot.snapshot(3)
    assert np.max(np.abs(a)) > 0
    assert np.max(np.abs(a - b)) <= 1e-3 * np.max(np.abs(a))


def test_layered_model_reduced_vs_oracle(gpu):
    """The layered anticline workload (reduced to 81 x 81 x 61, 300 steps), FP32 vs FP64 oracle."""
    cfg, model = workloads.layered(nx=81, ny=81, nz=61, nstep=300)
    cfg = cfg.replace(isrc=40, jsrc=40, ksrc=12)
    s32, e32 = run(cfg, model)
    s64, e64, _ = oracle.run(cfg, model, dtype=np.float64, threads=16)
    err = rel_l2(s32, s64)
    print(f"layered 81x81x61 FP32: seismogram rel L2 {err:.3e}")
    assert err <= 1e-4
    assert rel_l2(e32, e64) <= 1e-4
    s16, _ = run(cfg.replace(storage=STORAGE_FP16), model)
    print(f"layered 81x81x61 FP16 storage: seismogram rel L2 {rel_l2(s16, s64):.3e}")
    assert rel_l2(s16, s64) < 3e-2


def test_classic_shape_fp32_vs_gpu_fp64(gpu):
    """The 101 x 641 x 640 shape (600 steps): FP32 against the GPU FP64 instantiation,
    itself pinned to the CPU FP64 oracle by the small-grid parity tests."""
    cfg, model = workloads.classic(nstep=600)
    s32, e32 = run(cfg, model)
    s64, e64 = run(cfg.replace(storage=STORAGE_FP64), model)
    err = rel_l2(s32, s64)
    print(f"classic 101x641x640, 600 steps: FP32 vs FP64 seismogram rel L2 {err:.3e}")
    assert err <= 1e-4
    assert rel_l2(e32, e64) <= 1e-4
