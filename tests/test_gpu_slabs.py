"""GPU: a grid decomposed along z into slabs on one GPU, exchanging halos after
each half-step, equals the whole grid bit for bit.

This is the one-GPU part of a z decomposition: each slab owns a range of planes
plus a ghost plane on each side that has a neighbour, indexes its C-PML z tables
by global k (as the reference does at seismic_CPML_3D_isotropic_homogeneous_withCUDA.cu:177-179),
and receives through the pack/unpack kernels the planes the stencils read across
the interface (seismicCPML_3d_isotropic_simpelmodel.cpp:384, 450, 477 in the stress
pass; 508, 569 in the velocity pass).  Every cell does the same arithmetic in
the same order as in the whole grid, so fields, memory variables and receiver
samples are bit-identical; the energy is the same sum taken in another order.
"""
import numpy as np
import pytest

from ela3d_amd import (COMPAT_REFERENCE, SRC_EXPLOSIVE, SRC_FORCE3, STORAGE_FP32, STORAGE_FP64, Ela3dError,
                       Shot, set_tuning)
from helpers import hetero_model, small_cfg

pytestmark = pytest.mark.gpu


def run_whole(cfg, model):
    set_tuning(cfg, block=128, zseg_interior=2, vec=1, block_slab=128, zseg_slab=2, vec_slab=1)
    with Shot(cfg, *model) as s:
        s.run()
        return s.seismograms(), s.energy(), [s.snapshot(q) for q in range(27)]


def run_slabs(cfg, model, cuts):
    bounds = [0] + list(cuts) + [cfg.nz]
    slabs = [Shot(cfg, *model, slab=(a, b)) for a, b in zip(bounds[:-1], bounds[1:])]
    for _ in range(cfg.nstep):
        for phase in (0, 1):
            for s in slabs:
                s.step_phase(phase)
            for lower, upper in zip(slabs[:-1], slabs[1:]):
                lower.halo_exchange(upper, phase)
    return bounds, slabs


@pytest.mark.parametrize("storage,cuts,kind,compat,fs", [
    (STORAGE_FP32, (17,), SRC_EXPLOSIVE, 0, False),
    (STORAGE_FP64, (11, 23), SRC_FORCE3, 0, False),
    (STORAGE_FP32, (3, 16, 30), SRC_EXPLOSIVE, COMPAT_REFERENCE, False),  # cuts inside both z layers
    (STORAGE_FP32, (9, 20), SRC_FORCE3, 0, True),  # free surface on the lowest slab
    (STORAGE_FP64, (2, 17), SRC_EXPLOSIVE, COMPAT_REFERENCE, True),
])
def test_slabs_equal_the_whole_grid(gpu, storage, cuts, kind, compat, fs):
    cfg = small_cfg(nx=29, ny=23, nz=34, npml=4, compat=compat, storage=storage, nstep=45,
                    src_kind=kind, ksrc=cuts[0] - 1 if cuts[0] > 2 else 15, free_surface=fs)
    rec = np.asarray([(i, 11, k) for k in (0, 2, cuts[0] - 1, cuts[0], 20, 31) for i in range(1, 28, 5)],
                     dtype=np.int32)
    cfg = cfg.replace(receivers=rec)
    model = hetero_model(cfg, seed=9)
    print(f"slabs {cuts} storage {storage} source {kind} at k={cfg.ksrc} compat {compat} free surface {fs}",
          flush=True)
    seis, en, snaps = run_whole(cfg, model)
    bounds, slabs = run_slabs(cfg, model, cuts)
    try:
        got_seis = np.zeros_like(seis)
        got_en = np.zeros_like(en)
        for s, k0, k1 in zip(slabs, bounds[:-1], bounds[1:]):
            got_seis[:, s.receiver_index, :] = s.seismograms()
            got_en += s.energy()
            own = slice(k0 - s.koff, k1 - s.koff)
            for q in range(27):
                assert np.array_equal(s.snapshot(q)[own], snaps[q][k0:k1]), (q, k0, k1)
        assert sum(len(s.receiver_index) for s in slabs) == cfg.nrec
        assert np.any(seis) and np.array_equal(got_seis, seis)
        assert np.allclose(got_en, en, rtol=1e-12, atol=0.0)
    finally:
        for s in slabs:
            s.close()


def test_slab_halo_sides_and_sizes(gpu):
    cfg = small_cfg(nx=29, ny=23, nz=34, npml=4, nstep=2)
    model = hetero_model(cfg, seed=2)
    lower = Shot(cfg, *model, slab=(0, 17))
    upper = Shot(cfg, *model, slab=(17, 34))
    try:
        plane = 29 * 23 * 4
        assert lower.halo_bytes(0, 1) == 2 * plane and upper.halo_bytes(0, 0) == plane
        assert lower.halo_bytes(1, 1) == plane and upper.halo_bytes(1, 0) == 2 * plane
        with pytest.raises(Ela3dError):  # the lowest slab has no neighbour below
            lower.halo_exchange(lower, 0)
        with pytest.raises(Ela3dError):  # not neighbours in this order
            upper.halo_exchange(lower, 0)
    finally:
        lower.close()
        upper.close()
