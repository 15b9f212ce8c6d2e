"""Regression against tests/golden/small_fp64.npz, written by tests/golden/make_golden.py:
the FP64 oracle's seismograms and energy on two small deterministic runs (the clean
form and the reference's own choices).  The CPU test keeps the oracle from drifting;
the GPU test holds the FP64 HIP path to the same numbers."""
import importlib.util
import os

import numpy as np
import pytest

import oracle
from ela3d_amd import STORAGE_FP32, STORAGE_FP64, Shot

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
golden = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(golden)
DATA = np.load(os.path.join(HERE, "golden", "small_fp64.npz"))  # arrays and a JSON string only


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


@pytest.mark.parametrize("name", ["clean", "reference"])
def test_oracle_matches_golden(name):
    cfg = golden.configs()[name]
    seis, energy, _ = oracle.run(cfg, golden.model(), dtype=np.float64)
    assert np.abs(DATA[f"{name}_seis"]).max() > 0
    assert rel(seis, DATA[f"{name}_seis"]) <= 1e-12
    assert rel(energy, DATA[f"{name}_energy"]) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["clean", "reference"])
@pytest.mark.parametrize("storage", [STORAGE_FP64, STORAGE_FP32])
@pytest.mark.parametrize("fused", [False, True])
def test_gpu_matches_golden(gpu, name, storage, fused):
    """Both schedules.  The reference's own choices (a sigma source with the
    derivative-of-Gaussian wavelet) drive the stresses to about 1e17 Pa, so their
    squares and the energy terms built from them must stay within FP32."""
    from ela3d_amd import set_tuning
    cfg = golden.configs()[name].replace(storage=storage)
    set_tuning(cfg, block=128, zseg_interior=2, vec=1, block_slab=128, zseg_slab=2, vec_slab=1, fused=fused,
               tile=256)
    try:
        with Shot(cfg, *golden.model()) as shot:
            assert shot.info()["fused"] == int(fused)
            shot.run()
            seis, energy = shot.seismograms(), shot.energy()
    finally:
        set_tuning(cfg, block=128, zseg_interior=2, vec=1, block_slab=128, zseg_slab=2, vec_slab=1)
    tol = 1e-11 if storage == STORAGE_FP64 else 1e-5  # FP32: the workload's 1e-4 bar, tightened
    assert rel(seis, DATA[f"{name}_seis"]) <= tol
    assert rel(energy, DATA[f"{name}_energy"]) <= tol
