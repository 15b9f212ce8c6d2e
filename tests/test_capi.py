"""CPU checks of the C ABI: the library loads here (no GPU) and exports every
function include/ela3d.h declares; host-side validation and generators work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import ela3d_amd
from ela3d_amd import Config, Ela3dError
from ela3d_amd._lib import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(ela3d_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_config_default_and_validate():
    cfg = Config()
    cfg.validate()
    with pytest.raises(Ela3dError) as e:
        cfg.replace(nx=10, npoints_pml=10).validate()
    assert e.value.code == -1
    with pytest.raises(Ela3dError):
        cfg.replace(isrc=0).validate()
    with pytest.raises(Ela3dError):
        cfg.replace(receivers=np.array([[0, 0, 101]], np.int32)).validate()
    with pytest.raises(Ela3dError):
        cfg.replace(dt=-1.0).validate()


def test_padded_arrays_fit_32bit_byte_offsets():
    """Kernels address every field and model array by a 32-bit byte offset: a grid
    whose padded array would reach 4 GiB is refused, per storage element size."""
    from ela3d_amd import STORAGE_FP64
    big = Config().replace(nx=1024, ny=1024, nz=520, isrc=512, jsrc=512, ksrc=260)
    big.validate()  # FP32: 2.2 GB per padded array
    with pytest.raises(Ela3dError) as e:
        big.replace(storage=STORAGE_FP64).validate()  # FP64: 4.4 GB
    assert "4 GiB" in str(e.value)
    big.replace(nz=511, ksrc=255, storage=STORAGE_FP64).validate()  # 4.29e9 - 4 MB: still fits


def test_struct_size_guard():
    s, keep = Config().to_struct()
    s.struct_size = 12
    assert lib().ela3d_validate(C.byref(s)) == -1


def test_create_without_gpu_fails_loudly():
    if ela3d_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    cfg = Config(nx=23, ny=23, nz=23, npoints_pml=4, isrc=11, jsrc=11, ksrc=11)
    with pytest.raises(Ela3dError) as e:
        ela3d_amd.Shot(cfg, *ela3d_amd.homogeneous_model(23, 23, 23))
    assert e.value.code == -5


def test_homogeneous_model():
    cp, cs, rho = ela3d_amd.homogeneous_model(5, 6, 7, 3300.0, 1905.0, 2800.0)
    assert cp.shape == (7, 6, 5)
    assert np.all(cp == np.float32(3300.0)) and np.all(rho == np.float32(2800.0))


def test_layered_model_seeded_and_gardner():
    a = ela3d_amd.layered_model(41, 37, 31, 50.0, seed=7)
    b = ela3d_amd.layered_model(41, 37, 31, 50.0, seed=7)
    c = ela3d_amd.layered_model(41, 37, 31, 50.0, seed=8)
    assert all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3]))
    assert not np.array_equal(a[0], c[0])
    cp, cs, rho, vmax = a
    assert vmax == pytest.approx(float(cp.max()))
    # Vs = Vp/1.73 and rho = 310 Vp^0.25 per cell
    assert np.allclose(cs, cp / 1.73, rtol=2e-7)
    assert np.allclose(rho, 310.0 * cp.astype(np.float64) ** 0.25, rtol=1e-6)
    # six layer velocities 2000..4500 within +-3 %
    base = np.array([2000, 2500, 3000, 3500, 4000, 4500], np.float64)
    ratio = cp.astype(np.float64).ravel()[:, None] / base[None, :]
    assert np.all(np.min(np.abs(ratio - 1.0), axis=1) <= 0.0300001)
    # the anticline lifts layers at the centre: shallower top interface there
    col_c = cp[:, 18, 20]
    col_e = cp[:, 0, 0]
    first_c = np.argmax(col_c > 2000 * 1.04)
    first_e = np.argmax(col_e > 2000 * 1.04)
    assert first_c < first_e


def test_tuning_set_and_get_round_trip():
    """A pinned launch shape is kept per grid and read back field for field (no GPU
    needed); shapes outside the documented ranges are refused."""
    from ela3d_amd._lib import TuningStruct
    cfg = Config(nx=37, ny=41, nz=23)
    ela3d_amd.set_tuning(cfg, block=128, zseg_interior=12, vec=2, block_slab=64, zseg_slab=2, vec_slab=1,
                         fused=True, tile=512, strips=3)
    s, keep = cfg.to_struct()
    t = TuningStruct()
    assert lib().ela3d_tuning_get(C.byref(s), C.byref(t)) == 0
    assert (t.block, t.zseg_interior, t.vec, t.block_slab, t.zseg_slab, t.vec_slab) == (128, 12, 2, 64, 2, 1)
    assert (t.fused, t.dbuf, t.tile, t.strips) == (1, 1, 512, 3)
    for bad in (dict(strips=-1), dict(strips=65), dict(block=96), dict(vec=3)):
        kw = dict(block=128, zseg_interior=12, vec=2, block_slab=64, zseg_slab=2, vec_slab=1, fused=True,
                  tile=512, strips=3)
        kw.update(bad)
        with pytest.raises(Ela3dError):
            ela3d_amd.set_tuning(cfg, **kw)
