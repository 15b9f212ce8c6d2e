"""CPU: the algorithmic-traffic model of the two passes (bytes per step, split
between the interior launch and the absorbing-layer launch).

The byte count rests on which (memory variable, cell) pairs a pass updates.
The oracle determines that set independently: starting from zero memory
variables and random wavefields, one oracle step leaves a memory variable
non-zero exactly where its update runs with a != 0.  The library's plan must
then equal  cells * (bytes per cell of both passes) + 2 * element * count.
"""
import numpy as np
import pytest

import oracle
from ela3d_amd import STORAGE_FP16, STORAGE_FP32, STORAGE_FP64
from helpers import hetero_model, small_cfg


def touched(cfg, seed=4):
    """Per pass, the number of (memory variable, cell) pairs one step updates."""
    rng = np.random.default_rng(seed)
    st = oracle.State(cfg, np.float64)
    for q in range(9):
        st.f[q][...] = rng.standard_normal(st.f[q].shape)
    model = hetero_model(cfg, seed=seed)
    oracle.stress(cfg, model, st)
    oracle.velocity(cfg, model, st)
    stress = sum(int(np.count_nonzero(st.m[q])) for q in range(9))
    vel = sum(int(np.count_nonzero(st.m[q])) for q in range(9, 18))
    return stress, vel


@pytest.mark.parametrize("shape", [(29, 23, 26, 4), (40, 13, 15, 4), (17, 19, 18, 3), (64, 64, 9, 3)])
@pytest.mark.parametrize("compat", [0, 7])
@pytest.mark.parametrize("fs", [False, True])
def test_plan_bytes_match_oracle_update_set(shape, compat, fs):
    nx, ny, nz, npml = shape
    cfg = small_cfg(nx=nx, ny=ny, nz=nz, npml=npml, compat=compat, free_surface=fs)
    ns, nv = touched(cfg)
    cells = nx * ny * nz
    for storage, e in ((STORAGE_FP32, 4), (STORAGE_FP16, 2), (STORAGE_FP64, 8)):
        p = cfg.replace(storage=storage).plan()
        stress = (15 * e + 12) * cells + 2 * e * ns
        vel = (12 * e + 4) * cells + 2 * e * nv
        assert p["alg_bytes_stress"] == pytest.approx(stress, rel=1e-12)
        assert p["alg_bytes_velocity"] == pytest.approx(vel, rel=1e-12)
        assert p["alg_bytes_per_step"] == pytest.approx(stress + vel, rel=1e-12)
        assert p["pml_bytes_per_step"] == pytest.approx(2 * e * (ns + nv), rel=1e-9)


def test_classic_bytes_per_cell():
    from ela3d_amd import workloads
    cfg, _ = workloads.classic(nstep=10)
    p = cfg.plan()
    per_cell = p["alg_bytes_per_step"] / cfg.cells
    # 124 B/cell for the fields and model plus the C-PML share
    assert 124.0 < per_cell < 140.0
    assert p["pml_bytes_per_step"] == pytest.approx(p["alg_bytes_per_step"] - 124.0 * cfg.cells, rel=1e-9)
