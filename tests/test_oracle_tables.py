"""Pins the oracle's C-PML tables (CPU:142-361) and wavelets (CPU:608-610, UNI:441)
against closed forms, and checks the library's host-side tables against the oracle.

Closed forms used (Komatitsch & Martin 2007, Roden & Gedney 2000):
  d0 = -(N+1) vpml ln(R) / (2 L),  d(s) = d0 s^N,  alpha(s) = alpha_max (1 - s) + 0.1 alpha_max,
  alpha_max = pi f0;  the recursive-convolution coefficients satisfy
  b = exp(-(d/K + alpha) dt) and, at zero frequency, 1/K + a/(1-b) = 1/s(0) = alpha/(d + K alpha);
  the continuous layer reflects R = exp(-(2/vpml) int_0^L d dx).
"""
import math

import numpy as np
import pytest

import oracle
from ela3d_amd import (COMPAT_GHOST, COMPAT_REFERENCE, COMPAT_XHALF_A, COMPAT_ZHALF_MIN,
                       WAVELET_DGAUSSIAN, WAVELET_GAUSSIAN, WAVELET_RICKER, Config)


def cfg(**kw):
    base = dict(nx=101, ny=81, nz=61, dx=10.0, dy=10.0, dz=10.0, dt=1.6e-3, npoints_pml=10,
                f0=7.0, vpml=3000.0, rcoef=1e-3, npower=2.0, isrc=50, jsrc=40, ksrc=30)
    base.update(kw)
    return Config(**base)


def test_d0_and_alpha_max_spot_values():
    t = oracle.tables(cfg(), 0)
    # outermost full-grid point: s = 1 -> d = d0, alpha = 0.1 alpha_max
    d0 = 3.0 * 3000.0 * math.log(1000.0) / (2.0 * 100.0)
    assert d0 == pytest.approx(310.84899, abs=1e-4)
    assert t["d"][0] == pytest.approx(d0, rel=1e-14)
    assert t["d"][-1] == pytest.approx(d0, rel=1e-14)
    amax = math.pi * 7.0
    assert amax == pytest.approx(21.9911, abs=1e-4)
    assert t["alpha"][0] == pytest.approx(0.1 * amax, rel=1e-14)
    # the inner edge of the layer (s = 0) carries alpha = 1.1 alpha_max but d = 0
    assert t["d"][10] == 0.0 and t["alpha"][10] == pytest.approx(1.1 * amax, rel=1e-14)


@pytest.mark.parametrize("npower", [1.0, 2.0, 3.0])
def test_profile_power_law(npower):
    c = cfg(npower=npower)
    t = oracle.tables(c, 1)
    d0 = (npower + 1) * 3000.0 * math.log(1000.0) / 200.0
    for i in range(10):
        s_full = (10 - i) / 10.0
        s_half = (10 - i - 0.5) / 10.0
        assert t["d"][i] == pytest.approx(d0 * s_full ** npower, rel=1e-13)
        assert t["d_half"][i] == pytest.approx(d0 * s_half ** npower, rel=1e-13)


def test_mirror_symmetry_and_interior():
    for axis, n in ((0, 101), (1, 81), (2, 61)):
        t = oracle.tables(cfg(), axis)
        assert np.allclose(t["d"], t["d"][::-1], rtol=1e-12, atol=1e-9)
        # half point i sits at (i+1/2)h; its mirror is n-2-i
        assert np.allclose(t["d_half"][:n - 1], t["d_half"][:n - 1][::-1], rtol=1e-12, atol=1e-9)
        inner = slice(11, n - 11)
        for k in ("d", "d_half", "a", "a_half", "alpha", "alpha_half"):
            assert not np.any(t[k][inner]), (axis, k)
        assert np.all(t["b"][inner] == 1.0) and np.all(t["K"][inner] == 1.0)


@pytest.mark.parametrize("kmax", [1.0, 3.0])
def test_recursive_convolution_coefficients(kmax):
    c = cfg(k_max_pml=kmax)
    for axis in range(3):
        t = oracle.tables(c, axis)
        for sfx in ("", "_half"):
            d, K, al, a, b = (t[k + sfx] for k in ("d", "K", "alpha", "a", "b"))
            act = d > 1e-6
            assert act.sum() >= 10
            # b is the exact decay of exp(-(d/K + alpha) t) over one dt
            assert np.allclose(-np.log(b[act]) / c.dt, d[act] / K[act] + al[act], rtol=1e-12)
            # zero-frequency response of the stretched derivative: 1/s(0) = alpha/(d + K alpha)
            lhs = 1.0 / K[act] + a[act] / (1.0 - b[act])
            rhs = al[act] / (d[act] + K[act] * al[act])
            assert np.allclose(lhs, rhs, rtol=1e-9)


def test_discrete_reflection_coefficient():
    """Integrating the sampled profile (full and half points, spacing h/2) gives R."""
    c = cfg(nx=201, dx=5.0, npoints_pml=20)
    t = oracle.tables(c, 0)
    L = 20 * 5.0
    # samples at x = 0, h/2, h, ... across the min-side layer, abscissa L - x
    xs, ds = [], []
    for i in range(21):
        xs.append(i * 5.0)
        ds.append(t["d"][i])
        if i < 20:
            xs.append(i * 5.0 + 2.5)
            ds.append(t["d_half"][i])
    integral = np.trapezoid(ds, xs)
    R = math.exp(-2.0 * integral / c.vpml)
    assert R == pytest.approx(1e-3, rel=0.02)


def test_compat_quirks():
    clean = [oracle.tables(cfg(dz=8.0), a) for a in range(3)]
    q1 = oracle.tables(cfg(dz=8.0, compat_flags=COMPAT_XHALF_A), 0)
    assert not np.any(q1["a_half"]) and np.any(clean[0]["a_half"])
    assert np.array_equal(q1["a"], clean[0]["a"]) and np.array_equal(q1["b_half"], clean[0]["b_half"])
    q2 = oracle.tables(cfg(dz=8.0, compat_flags=COMPAT_ZHALF_MIN), 2)
    n = 61
    # no z-half absorption on the min side (stale y abscissa is far outside the layer)
    assert not np.any(q2["d_half"][: n // 2]) and np.any(clean[2]["d_half"][: n // 2])
    # the other sides normalise by the y thickness (100 m) instead of the z thickness (80 m)
    d0z = 3.0 * 3000.0 * math.log(1000.0) / (2.0 * 80.0)
    assert q2["d"][0] == pytest.approx(d0z * (80.0 / 100.0) ** 2, rel=1e-12)
    assert q2["d"][-1] == pytest.approx(d0z, rel=1e-12)       # max side full: z thickness
    # the ghost quirk changes no table
    g = oracle.tables(cfg(dz=8.0, compat_flags=COMPAT_GHOST), 2)
    assert all(np.array_equal(g[k], clean[2][k]) for k in g)


def test_no_pml():
    t = oracle.tables(cfg(npoints_pml=0), 0)
    assert not np.any(t["a"]) and not np.any(t["a_half"]) and np.all(t["b"] == 1.0)


@pytest.mark.parametrize("compat", [0, COMPAT_REFERENCE])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_library_tables_match_oracle(axis, compat):
    """The product's host-side FP64 table builder agrees with the oracle bit for bit."""
    c = cfg(compat_flags=compat, dy=12.0, dz=8.0, k_max_pml=2.0)
    lib_t = c.pml_tables(axis)
    orc_t = oracle.tables(c, axis)
    for k in orc_t:
        assert np.array_equal(lib_t[k], orc_t[k]), k


def test_plan_regions_cover_every_active_coefficient():
    for compat in (0, COMPAT_REFERENCE):
        c = cfg(compat_flags=compat)
        p = c.plan()
        for axis in range(3):
            t = oracle.tables(c, axis)
            lo, hi = p["region_lo"][axis], p["region_hi"][axis]
            act = np.nonzero((t["a"] != 0) | (t["a_half"] != 0))[0]
            assert np.all((act < lo) | (act >= hi))
            assert lo <= 10 and hi >= (101, 81, 61)[axis] - 11


def test_wavelet_closed_forms():
    c = cfg(f0=7.0, dt=1e-4)
    n = 4000
    t = np.arange(n) * c.dt
    a = math.pi ** 2 * 49.0
    g = oracle.wavelet(c.replace(wavelet=WAVELET_GAUSSIAN), n)
    dg = oracle.wavelet(c.replace(wavelet=WAVELET_DGAUSSIAN), n)
    rk = oracle.wavelet(c.replace(wavelet=WAVELET_RICKER), n)
    t0 = 1.2 / 7.0
    # first sample at t = 0 (CPU:609: t = (it-1) dt)
    assert g[0] == pytest.approx(1e7 * math.exp(-a * t0 * t0), rel=1e-14)
    assert g[int(round(t0 / c.dt))] == pytest.approx(1e7, rel=1e-6)
    # dGaussian is d/dt of the Gaussian; Ricker is -1/(2a) d2/dt2 of the Gaussian
    d1 = np.gradient(g, c.dt)
    assert np.max(np.abs(d1[5:-5] - dg[5:-5])) <= 2e-3 * np.max(np.abs(dg))
    d2 = np.gradient(d1, c.dt)
    assert np.max(np.abs(-d2[10:-10] / (2 * a) - rk[10:-10])) <= 5e-3 * np.max(np.abs(rk))
    # the Ricker spectrum peaks at f0
    spec = np.abs(np.fft.rfft(rk, 1 << 16))
    f = np.fft.rfftfreq(1 << 16, c.dt)
    assert f[np.argmax(spec)] == pytest.approx(7.0, abs=0.05)
