"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the same inputs.

Bars: FP32 storage vs the FP32 oracle element-wise within 2e-6 of each field's
max after one step from a random state (arithmetic is FP32 on both sides; the
GPU contracts to FMA, the oracle does not); FP64 storage vs the FP64 oracle
within 1e-12; seismograms of FP32 runs vs the FP64 oracle within relative
L2 1e-4 (the workload's stated tolerance).
"""
import numpy as np
import pytest

import oracle
from ela3d_amd import (COMPAT_REFERENCE, SRC_EXPLOSIVE, SRC_FORCE, SRC_FORCE3,
                       SRC_REFERENCE_SIGMA, STORAGE_FP32, STORAGE_FP64, WAVELET_DGAUSSIAN,
                       WAVELET_GAUSSIAN, WAVELET_RICKER, Shot)
from ela3d_amd import workloads
from helpers import (compare_state, hetero_model, homo, max_rel, push_state, random_state,
                     rel_l2, small_cfg)

pytestmark = pytest.mark.gpu

TOL = {STORAGE_FP32: 2e-6, STORAGE_FP64: 1e-12}
DT = {STORAGE_FP32: np.float32, STORAGE_FP64: np.float64}


def _one_step(cfg, model, seed):
    st = random_state(cfg, seed=seed, dtype=DT[cfg.storage])
    with Shot(cfg, *model) as shot:
        push_state(shot, st)
        shot.step(1)
        shot.sync()
        seis_o, en_o, st = oracle.run(cfg, model, 1, st=st, dtype=DT[cfg.storage])
        worst = compare_state(shot, st, TOL[cfg.storage])
        seis_g = shot.seismograms()[:1]
        en_g = shot.energy()[:1]
    assert max_rel(seis_g, seis_o[:1]) <= TOL[cfg.storage]
    # energy sums squares of the state, so it inherits the state's relative error
    assert np.allclose(en_g, en_o[:1], rtol=10 * TOL[cfg.storage], atol=0)
    return worst


@pytest.mark.parametrize("storage", [STORAGE_FP32, STORAGE_FP64])
@pytest.mark.parametrize("compat", [0, COMPAT_REFERENCE])
def test_one_step_random_state(gpu, storage, compat):
    cfg = small_cfg(compat=compat, storage=storage)
    print(f"one-step parity compat={compat} storage={storage}", flush=True)
    _one_step(cfg, hetero_model(cfg, seed=3), seed=11)


@pytest.mark.parametrize("src", [SRC_REFERENCE_SIGMA, SRC_FORCE, SRC_EXPLOSIVE, SRC_FORCE3])
def test_one_step_sources(gpu, src):
    cfg = small_cfg(src_kind=src, angle_force=30.0)
    _one_step(cfg, hetero_model(cfg, seed=4), seed=5)


@pytest.mark.parametrize("shape", [(17, 19, 18, 3), (40, 13, 15, 4), (11, 12, 13, 0), (64, 64, 9, 3)])
def test_one_step_ragged_shapes(gpu, shape):
    nx, ny, nz, npml = shape
    cfg = small_cfg(nx=nx, ny=ny, nz=nz, npml=npml)
    _one_step(cfg, hetero_model(cfg, seed=7), seed=8)


@pytest.mark.parametrize("wavelet", [WAVELET_GAUSSIAN, WAVELET_DGAUSSIAN, WAVELET_RICKER])
@pytest.mark.parametrize("compat", [0, COMPAT_REFERENCE])
def test_from_rest_seismograms(gpu, wavelet, compat):
    cfg = small_cfg(compat=compat, wavelet=wavelet, nstep=120)
    model = hetero_model(cfg, seed=9)
    with Shot(cfg, *model) as shot:
        shot.run()
        sg, eg = shot.seismograms(), shot.energy()
    s32, e32, _ = oracle.run(cfg, model, dtype=np.float32)
    s64, e64, _ = oracle.run(cfg, model, dtype=np.float64)
    assert rel_l2(sg, s32) <= 2e-5
    assert rel_l2(sg, s64) <= 1e-4
    assert rel_l2(eg, e64) <= 1e-4


def test_fp64_storage_matches_fp64_oracle(gpu):
    cfg = small_cfg(storage=STORAGE_FP64, nstep=80)
    model = hetero_model(cfg, seed=2)
    with Shot(cfg, *model) as shot:
        shot.run()
        sg, eg = shot.seismograms(), shot.energy()
    s64, e64, _ = oracle.run(cfg, model, dtype=np.float64)
    assert rel_l2(sg, s64) <= 1e-11
    assert rel_l2(eg, e64) <= 1e-11


def test_graph_and_eager_identical(gpu):
    cfg = small_cfg(nstep=50)
    model = homo(cfg)
    out = []
    for g in (True, False):
        with Shot(cfg.replace(use_graph=g), *model) as shot:
            shot.run()
            out.append((shot.seismograms(), shot.energy(), shot.snapshot(1)))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


SHAPES = [  # interior launch, absorbing-layer launch, schedule
    dict(block=256, zseg_interior=4, vec=1, block_slab=256, zseg_slab=4, vec_slab=1),
    dict(block=256, zseg_interior=4, vec=1, block_slab=256, zseg_slab=4, vec_slab=1, overlap=True),
    dict(block=256, zseg_interior=4, vec=1, block_slab=256, zseg_slab=4, vec_slab=1, dbuf=True),
    dict(block=64, zseg_interior=2, vec=2, block_slab=128, zseg_slab=2, vec_slab=2, overlap=True, dbuf=True),
    dict(block=128, zseg_interior=8, vec=4, block_slab=64, zseg_slab=8, vec_slab=1),
    dict(block=64, zseg_interior=32, vec=4, block_slab=256, zseg_slab=2, vec_slab=2, overlap=True),
    dict(block=256, zseg_interior=4, vec=4, block_slab=256, zseg_slab=2, vec_slab=1, merged=True),
    dict(block=128, zseg_interior=2, vec=2, block_slab=128, zseg_slab=8, vec_slab=1, merged=True, dbuf=True),
    dict(block=256, zseg_interior=8, vec=1, block_slab=256, zseg_slab=4, vec_slab=1, fused=True, tile=512),
    dict(block=256, zseg_interior=3, vec=1, block_slab=128, zseg_slab=2, vec_slab=2, fused=True, tile=256),
    dict(block=256, zseg_interior=32, vec=1, block_slab=256, zseg_slab=4, vec_slab=1, fused=True, tile=1024),
    dict(block=256, zseg_interior=4, vec=1, block_slab=128, zseg_slab=2, vec_slab=1, fused=True, tile=256, strips=3),
]


@pytest.mark.parametrize("storage", [STORAGE_FP32, STORAGE_FP64])
def test_launch_shapes_agree(gpu, storage):
    """Every launch shape and schedule computes the same cells: the first three
    (overlapped launches, double-buffered field sets) are bit-identical; the others
    agree to rounding (thread counts, z segments, VEC widths, the mixed launch and
    the fused interior step change only how the work is cut; energy partial sums
    are grouped by workgroup)."""
    from ela3d_amd import set_tuning
    cfg = small_cfg(nx=45, ny=38, nz=33, nstep=60, storage=storage)
    model = hetero_model(cfg, seed=5)
    out = []
    for sh in SHAPES:
        if storage == STORAGE_FP64 and sh.get("tile", 0) > 512:  # FP64 LDS planes fit 512-cell tiles
            sh = dict(sh, tile=512)
        set_tuning(cfg, **sh)
        with Shot(cfg, *model) as shot:
            inf = shot.info()
            assert inf["fused"] == int(sh.get("fused", False))
            assert inf["dbuf"] == int(sh.get("dbuf", False) or sh.get("fused", False))
            assert inf["merged"] == int(sh.get("merged", False))
            assert inf["strips"] == (sh.get("strips", 0) if sh.get("fused") else 0)
            if not sh.get("fused"):
                assert (inf["block"], inf["vec"], inf["vec_slab"]) == (sh["block"], sh["vec"], sh["vec_slab"])
            shot.run()
            out.append((shot.seismograms(), shot.energy(), shot.snapshot(5), shot.snapshot(20),
                        shot.snapshot(1), shot.snapshot(24)))
    for o in out[1:3]:
        for x, y in zip(out[0], o):
            assert np.array_equal(x, y)
    tol = 1e-6 if storage == STORAGE_FP32 else 1e-13
    for n, o in enumerate(out[3:], start=3):
        exact = all(np.array_equal(x, y) for x, y in zip(out[0], o))
        print(f"shape {n} {SHAPES[n]}: {'bit-identical' if exact else 'rounding-level'}; "
              f"seismogram rel L2 {rel_l2(o[0], out[0][0]):.2e}")
        for q in (0, 2, 3, 4, 5):
            assert rel_l2(o[q], out[0][q]) <= tol, (n, q)
        assert np.allclose(o[1], out[0][1], rtol=1e-12 if storage == STORAGE_FP64 else 1e-6)
    set_tuning(cfg, **SHAPES[0])


@pytest.mark.parametrize("storage,tile", [(STORAGE_FP32, 768), (STORAGE_FP64, 512)])
def test_fused_strip_order_default(gpu, storage, tile):
    """A grid with enough tiles per plane (at least 32) that the fused step walks its
    default 4-strip order: the same cells as the two-pass schedule, to rounding."""
    from ela3d_amd import set_tuning
    cfg = small_cfg(nx=48, ny=600, nz=30, nstep=30, storage=storage)
    model = hetero_model(cfg, seed=8)
    out = []
    for fused in (False, True):
        set_tuning(cfg, block=128, zseg_interior=8, vec=1, block_slab=128, zseg_slab=2, vec_slab=1, fused=fused,
                   tile=tile)
        with Shot(cfg, *model) as shot:
            inf = shot.info()
            assert inf["fused"] == int(fused) and inf["strips"] == (4 if fused else 0)
            shot.run()
            out.append((shot.seismograms(), shot.snapshot(0), shot.snapshot(5), shot.snapshot(8)))
    set_tuning(cfg, block=128, zseg_interior=8, vec=1, block_slab=128, zseg_slab=2, vec_slab=1)
    tol = 1e-6 if storage == STORAGE_FP32 else 1e-13
    assert np.abs(out[0][0]).max() > 0
    for x, y in zip(out[0], out[1]):
        assert rel_l2(y, x) <= tol


@pytest.mark.parametrize("storage", [STORAGE_FP32, STORAGE_FP64])
@pytest.mark.parametrize("tile,zseg", [(512, 4), (256, 7), (1024, 1)])
def test_fused_one_step_random_state(gpu, storage, tile, zseg):
    """The fused schedule against the oracle after one step from a random state
    (every field and memory variable non-zero where it may be)."""
    from ela3d_amd import set_tuning
    cfg = small_cfg(nx=37, ny=29, nz=27, storage=storage, nstep=3)
    if storage == STORAGE_FP64:
        tile = min(tile, 512)  # FP64 fused workgroups hold at most 512 threads
    set_tuning(cfg, block=256, zseg_interior=zseg, zseg_slab=2, fused=True, tile=tile)
    model = hetero_model(cfg, seed=7)
    with Shot(cfg, *model) as shot:
        assert shot.info()["fused"] == 1 and shot.info()["tile"] == tile
    worst, name = _one_step(cfg, model, seed=11)
    print(f"fused tile {tile} zseg {zseg}: worst {worst:.2e} ({name})")


def test_zero_source_stays_zero(gpu):
    cfg = small_cfg(factor=0.0, nstep=20)
    with Shot(cfg, *homo(cfg)) as shot:
        shot.run()
        for q in range(27):
            assert not np.any(shot.snapshot(q)), q


@pytest.mark.parametrize("compat", [0, COMPAT_REFERENCE])
def test_homogeneous_cube_vs_fp64_oracle(gpu, compat):
    """101^3, 500 steps: FP32 seismograms within relative L2 1e-4 of the FP64 oracle."""
    cfg, model = workloads.homogeneous_cube(compat_flags=compat, dirichlet=(compat == 0))
    print(f"homogeneous cube compat={compat}", flush=True)
    with Shot(cfg, *model) as shot:
        shot.run()
        sg, eg = shot.seismograms(), shot.energy()
    s64, e64, _ = oracle.run(cfg, model, dtype=np.float64, threads=16)
    err = rel_l2(sg, s64)
    print(f"seismogram rel L2 vs FP64 oracle: {err:.3e}", flush=True)
    assert err <= 1e-4
    assert rel_l2(eg, e64) <= 1e-4
    # C-PML stability: after the peak the interior energy never exceeds it, and
    # with the intended (clean) tables the wave leaves the box through the layers
    tot = eg.sum(axis=1)
    peak = int(np.argmax(tot))
    assert np.all(tot[peak:] <= tot[peak])
    if compat == 0:
        assert tot[-1] < 0.02 * tot[peak]


def test_empty_receivers_energy_off_and_zero_steps(gpu):
    """Edge cases of the boundary.  An empty receiver list with the energy off runs,
    returns empty outputs and matches the oracle's fields after its steps; a
    config with nstep = 0 creates, runs nothing and returns empty outputs."""
    cfg = small_cfg(receivers=np.zeros((0, 3), np.int32), energy_every=0, nstep=3)
    model = hetero_model(cfg, seed=9)
    st = random_state(cfg, seed=10)
    with Shot(cfg, *model) as shot:
        push_state(shot, st)
        shot.run()
        shot.sync()
        inf = shot.info()
        assert inf["steps_done"] == 3 and inf["nenergy"] == 0
        assert shot.seismograms().shape == (inf["nsamples"], 0, 3)
        assert shot.energy().shape == (0, 2)
        _, _, st = oracle.run(cfg, model, 3, st=st)
        compare_state(shot, st, TOL[STORAGE_FP32])
    cfg0 = small_cfg(nstep=0)
    with Shot(cfg0, *homo(cfg0)) as shot:
        shot.run()
        shot.sync()
        assert shot.info()["steps_done"] == 0
        assert shot.seismograms().size == 0 and shot.energy().size == 0


def test_sampled_profile_same_results_as_plain_steps(gpu):
    """profile(n, every) interleaves instrumented eager steps (hipEvents around each
    launch) with graph replays: the fields, seismograms and energy equal n plain
    steps bit for bit, and the per-launch totals cover exactly the instrumented steps."""
    cfg = small_cfg(nstep=24)
    model = hetero_model(cfg, seed=12)
    with Shot(cfg, *model) as a, Shot(cfg, *model) as b:
        a.step(23)
        a.sync()
        ks, wall = b.profile(23, every=5)  # steps 0, 5, 10, 15, 20 instrumented
        b.sync()
        assert wall > 0 and b.info()["steps_done"] == 23
        assert [k["launches"] for k in ks] == [5] * len(ks)
        for q in range(27):
            np.testing.assert_array_equal(a.snapshot(q), b.snapshot(q))
        np.testing.assert_array_equal(a.seismograms(), b.seismograms())
        np.testing.assert_array_equal(a.energy(), b.energy())


def test_autotune_shape_is_used_and_correct(gpu):
    """The autotuner's pick is cached for the grid, used by the next shot, and that
    shot meets the seismogram bar against the FP64 oracle.  Every candidate shape the
    tuner tries on this grid (split, merged and fused) passes the CPU address model."""
    from ela3d_amd import autotune
    cfg = small_cfg(nx=45, ny=38, nz=33, nstep=80)
    model = hetero_model(cfg, seed=6)
    t = autotune(cfg, model, steps=2)
    print("autotune:", t)
    assert t["block"] in (64, 128, 256) and t["vec"] in (1, 2, 4) and t["zseg_interior"] in (1, 2, 4, 8, 16, 32)
    assert t["candidates"] > 50 and t["ms_per_step"] > 0
    with Shot(cfg, *model) as shot:
        inf = shot.info()
        assert inf["fused"] == t["fused"] and inf["merged"] == t["merged"]
        if not t["fused"]:
            assert (inf["block"], inf["vec"]) == (t["block"], t["vec"])
        shot.run()
        seis = shot.seismograms()
    s64, _, _ = oracle.run(cfg, model, dtype=np.float64)
    assert rel_l2(seis, s64) <= 1e-4
