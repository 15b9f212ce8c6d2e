"""GPU: checkpoint/resume, the source trace, and the ela3d_run CLI end to end.

A resumed shot must continue bit-identically (the checkpoint is the whole device
state); the source trace is checked against the FP32 oracle's stresses at the
source after every step; the CLI's output files must hold exactly what the C ABI
returns for the same run.
"""
import os
import subprocess

import numpy as np
import pytest

import ela3d_amd
import oracle
from ela3d_amd import SRC_EXPLOSIVE, SRC_REFERENCE_SIGMA, STORAGE_FP16, Shot
from helpers import hetero_model, homo, small_cfg

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(ela3d_amd.__file__), "bin", "ela3d_run")


@pytest.mark.parametrize("storage,dbuf,fs", [(0, False, False), (STORAGE_FP16, False, False), (0, True, False),
                                           (0, True, True)])
def test_checkpoint_resume_bit_identical(gpu, tmp_path, storage, dbuf, fs):
    """Also with two field sets (the set holding the state follows the step count) and
    with a free surface."""
    cfg = small_cfg(nx=31 if dbuf else 29, nstep=90, storage=storage, src_kind=SRC_EXPLOSIVE, free_surface=fs)
    if dbuf:
        ela3d_amd.set_tuning(cfg, block=256, zseg_interior=4, zseg_slab=4, dbuf=True)
    model = hetero_model(cfg, seed=2)
    with Shot(cfg, *model) as a:
        a.run()
        ref = (a.seismograms(), a.energy(), a.snapshot(3), a.snapshot(22), a.source_trace())
    ck = str(tmp_path / "shot.ck")
    with Shot(cfg, *model) as b:
        b.step(37)
        b.sync()
        b.checkpoint(ck)
        b.step(5)  # diverge after the checkpoint; the restore must undo it
    with Shot(cfg, *model) as c:
        c.restore(ck)
        assert c.info()["steps_done"] == 37
        c.run()
        got = (c.seismograms(), c.energy(), c.snapshot(3), c.snapshot(22), c.source_trace())
    for x, y in zip(ref, got):
        assert np.array_equal(x, y)


def test_checkpoint_rejects_other_config(gpu, tmp_path):
    cfg = small_cfg(nstep=20)
    ck = str(tmp_path / "a.ck")
    with Shot(cfg, *homo(cfg)) as a:
        a.step(3)
        a.checkpoint(ck)
    other = cfg.replace(f0=cfg.f0 * 1.5)
    with Shot(other, *homo(other)) as b:
        with pytest.raises(ela3d_amd.Ela3dError) as e:
            b.restore(ck)
        assert e.value.code == -1
    surface = cfg.replace(free_surface=True)  # same sizes, another boundary: refused too
    with Shot(surface, *homo(surface)) as b:
        with pytest.raises(ela3d_amd.Ela3dError) as e:
            b.restore(ck)
        assert e.value.code == -1
        with pytest.raises(ela3d_amd.Ela3dError) as e:
            b.restore(str(tmp_path / "missing.ck"))
        assert e.value.code == -11


def test_checkpoint_truncated_leaves_shot_unchanged(gpu, tmp_path):
    cfg = small_cfg(nstep=20)
    ck = str(tmp_path / "a.ck")
    with Shot(cfg, *homo(cfg)) as a:
        a.step(9)
        a.checkpoint(ck)
    with open(ck, "rb") as f:
        whole = f.read()
    short = str(tmp_path / "short.ck")
    with open(short, "wb") as f:
        f.write(whole[:-4096])
    with Shot(cfg, *homo(cfg)) as b:
        b.step(4)
        before = [b.snapshot(q) for q in (0, 3)]
        with pytest.raises(ela3d_amd.Ela3dError) as e:
            b.restore(short)
        assert e.value.code == -11
        assert b.info()["steps_done"] == 4
        for q, x in zip((0, 3), before):
            assert np.array_equal(b.snapshot(q), x)


@pytest.mark.parametrize("src", [SRC_EXPLOSIVE, SRC_REFERENCE_SIGMA])
def test_source_trace_vs_oracle(gpu, src):
    cfg = small_cfg(nstep=30, src_kind=src)
    model = hetero_model(cfg, seed=9)
    with Shot(cfg, *model) as s:
        s.run()
        tr = s.source_trace()
    st = oracle.State(cfg, np.float32)
    want = np.zeros_like(tr)
    o = (cfg.ksrc, cfg.jsrc, cfg.isrc)
    for it in range(cfg.nstep):
        oracle.run(cfg, model, 1, st=st, it0=it, dtype=np.float32)
        want[it + 1] = [st.f[3][o], st.f[4][o], st.f[5][o]]
    assert not np.any(tr[0])
    scale = np.abs(want).max()
    assert scale > 0
    assert np.abs(tr - want).max() <= 2e-6 * scale


@pytest.mark.skipif(not os.path.exists(CLI), reason="ela3d_run not built")
def test_cli_free_surface(gpu, tmp_path):
    """free_surface=1 on the command line runs the library's free surface."""
    n, npml, nstep = 28, 4, 30
    args = [f"nx={n}", f"ny={n}", f"nz={n}", f"npml={npml}", f"nstep={nstep}", "dt=1e-3", "f0=25",
            "rec=line:x,14,0,4", "free_surface=1", f"out={tmp_path}"]
    r = subprocess.run([CLI, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = np.array([(i, 14, 0) for i in range(0, n, 4)], np.int32)
    cfg = ela3d_amd.Config(nx=n, ny=n, nz=n, npoints_pml=npml, nstep=nstep, dt=1e-3, f0=25.0,
                           isrc=n // 2, jsrc=n // 2, ksrc=n // 2, receivers=rec, free_surface=True)
    model = ela3d_amd.homogeneous_model(n, n, n, 3300.0, 3300.0 / 1.732, 2800.0)
    with Shot(cfg, *model) as s:
        s.run()
        seis = s.seismograms()
    got = np.array([[float(x) for x in ln.split()] for ln in open(tmp_path / "rechorvx.txt").read().splitlines()])
    assert got.shape == seis[:, :, 0].shape
    assert np.allclose(got, seis[:, :, 0].astype(np.float32).astype(np.float64), rtol=5e-6, atol=1e-30)
    assert np.any(seis[:, :, 0] != 0.0)  # the surface moves (without it, vx on plane 0 stays 0)


def test_cli_end_to_end(gpu, tmp_path):
    n, npml, nstep = 32, 4, 40
    args = [f"nx={n}", f"ny={n}", f"nz={n}", f"npml={npml}", f"nstep={nstep}", "dt=1e-3", "f0=25",
            "rec=grid:8,4,6", "sampgat=2", "it_output=20", "energy_every=5", f"out={tmp_path}"]
    r = subprocess.run([CLI, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # the same run through the library
    xlen, ylen = n // 8, (n - 2 * npml) // 4
    rec = np.array([(i, j, 6) for j in range(0, n, ylen) for i in range(0, n, xlen)], np.int32)
    cfg = ela3d_amd.Config(nx=n, ny=n, nz=n, npoints_pml=npml, nstep=nstep, dt=1e-3, f0=25.0,
                           isrc=n // 2, jsrc=n // 2, ksrc=n // 2, receivers=rec, sample_every=2,
                           energy_every=5)
    model = ela3d_amd.homogeneous_model(n, n, n, 3300.0, 3300.0 / 1.732, 2800.0)
    with Shot(cfg, *model) as s:
        s.step(20)
        s.sync()
        vz20 = s.snapshot(2).astype(np.float32)
        s.run()
        seis, en, tr = s.seismograms(), s.energy(), s.source_trace()
        vx40 = s.snapshot(0).astype(np.float32)
    for comp, name in enumerate(("rechorvx.txt", "rechorvy.txt", "rechorvz.txt")):
        lines = open(tmp_path / name).read().splitlines()
        assert len(lines) == nstep // 2
        got = np.array([[float(x) for x in ln.split()] for ln in lines])
        want = seis[:, :, comp].astype(np.float32).astype(np.float64)
        assert np.allclose(got, want, rtol=5e-6, atol=1e-30)
    assert np.array_equal(np.fromfile(tmp_path / "vz00020.bin", "<f4"), vz20.ravel())
    assert np.array_equal(np.fromfile(tmp_path / "vx00040.bin", "<f4"), vx40.ravel())
    e = np.loadtxt(tmp_path / "energy.txt")
    assert np.array_equal(e[:, 0], np.arange(5, nstep + 1, 5))
    assert np.allclose(e[:, 1:], en, rtol=1e-15, atol=0)
    src = open(tmp_path / "recsrcz.txt").read()
    assert not src.endswith("\n") and len(src.split()) == nstep + 1
