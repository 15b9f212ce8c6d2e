"""GPU: the free surface (traction-free plane k = 0, no z-min C-PML) against the
oracle, whose free surface is pinned by exact discrete solutions, the plane-stress
law and the Rayleigh-wave speed (test_oracle_free_surface.py).

Bars as for the other parity tests: after one step from a random state, FP32
storage element-wise within 2e-6 of each field's maximum against the FP32 oracle
and FP64 within 1e-12; seismograms from rest within relative L2 1e-4 of the FP64
oracle (FP32) or 3e-2 (FP16 storage); z-slabs bit-identical to the whole grid.
"""
import numpy as np
import pytest

import oracle
from ela3d_amd import (COMPAT_REFERENCE, SRC_EXPLOSIVE, SRC_FORCE3, STORAGE_FP16, STORAGE_FP32,
                       STORAGE_FP64, WAVELET_RICKER, Shot, set_tuning)
from helpers import compare_state, hetero_model, max_rel, push_state, random_state, rel_l2, small_cfg

pytestmark = pytest.mark.gpu

TOL = {STORAGE_FP32: 2e-6, STORAGE_FP64: 1e-12}
DT = {STORAGE_FP32: np.float32, STORAGE_FP64: np.float64}

SCHEDULES = [
    dict(block=256, zseg_interior=4, vec=1, block_slab=256, zseg_slab=4, vec_slab=1),
    dict(block=64, zseg_interior=2, vec=2, block_slab=128, zseg_slab=1, vec_slab=2, dbuf=True),
    dict(block=128, zseg_interior=8, vec=4, block_slab=64, zseg_slab=3, vec_slab=1, merged=True),
    dict(block=256, zseg_interior=5, vec=1, block_slab=128, zseg_slab=2, vec_slab=1, fused=True, tile=512),
]


def fs_cfg(**kw):
    base = dict(free_surface=True, ksrc=3)
    base.update(kw)
    return small_cfg(**base)


@pytest.mark.parametrize("storage", [STORAGE_FP32, STORAGE_FP64])
@pytest.mark.parametrize("compat", [0, COMPAT_REFERENCE])
@pytest.mark.parametrize("sched", range(len(SCHEDULES)))
def test_one_step_random_state(gpu, storage, compat, sched):
    cfg = fs_cfg(compat=compat, storage=storage, nstep=3)
    set_tuning(cfg, **SCHEDULES[sched])
    model = hetero_model(cfg, seed=3)
    st = random_state(cfg, seed=11, dtype=DT[storage])
    print(f"free surface one step: storage {storage} compat {compat} schedule {SCHEDULES[sched]}", flush=True)
    with Shot(cfg, *model) as shot:
        assert shot.info()["fused"] == int(SCHEDULES[sched].get("fused", False))
        push_state(shot, st)
        shot.step(1)
        shot.sync()
        seis_o, _, st = oracle.run(cfg, model, 1, st=st, dtype=DT[storage])
        worst, name = compare_state(shot, st, TOL[storage])
        assert max_rel(shot.seismograms()[:1], seis_o[:1]) <= TOL[storage]
    print(f"  worst {worst:.2e} ({name})")


@pytest.mark.parametrize("shape", [(11, 12, 13, 0), (40, 13, 15, 4), (17, 19, 18, 3)])
def test_one_step_ragged_shapes(gpu, shape):
    nx, ny, nz, npml = shape
    cfg = fs_cfg(nx=nx, ny=ny, nz=nz, npml=npml, nstep=3)
    model = hetero_model(cfg, seed=7)
    st = random_state(cfg, seed=8, dtype=np.float32)
    with Shot(cfg, *model) as shot:
        push_state(shot, st)
        shot.step(1)
        shot.sync()
        oracle.run(cfg, model, 1, st=st, dtype=np.float32)
        compare_state(shot, st, TOL[STORAGE_FP32])


def surface_shot(storage=STORAGE_FP32, nstep=220):
    rec = np.asarray([(i, 20, 0) for i in range(6, 40, 3)] + [(20, j, 1) for j in range(6, 40, 4)], np.int32)
    cfg = small_cfg(nx=46, ny=44, nz=30, npml=6, free_surface=True, src_kind=SRC_FORCE3, ksrc=2, isrc=20,
                    jsrc=20, receivers=rec, nstep=nstep, storage=storage, f0=14.0, dirichlet=False)
    return cfg, hetero_model(cfg, seed=12)


def test_from_rest_against_fp64_oracle(gpu):
    cfg, model = surface_shot()
    seis_o, en_o, _ = oracle.run(cfg, model, dtype=np.float64, threads=8)
    with Shot(cfg, *model) as shot:
        shot.run()
        seis, en = shot.seismograms(), shot.energy()
    err = rel_l2(seis, seis_o)
    print(f"free surface from rest: seismogram rel L2 vs FP64 oracle {err:.2e}")
    assert err <= 1e-4
    assert np.allclose(en, en_o, rtol=1e-4)


def test_fp16_storage_from_rest(gpu):
    cfg, model = surface_shot(storage=STORAGE_FP16)
    seis_o, _, _ = oracle.run(cfg.replace(storage=STORAGE_FP64), model, dtype=np.float64, threads=8)
    with Shot(cfg, *model) as shot:
        shot.run()
        err = rel_l2(shot.seismograms(), seis_o)
    print(f"free surface FP16 storage: seismogram rel L2 vs FP64 oracle {err:.2e}")
    assert err <= 3e-2
