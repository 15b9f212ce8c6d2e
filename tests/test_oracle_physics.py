"""Pins the oracle's stencils (CPU:377-590), source (CPU:604-622), receivers
(CPU:593-602) and energy against what the mathematics fixes.

* Discrete plane waves.  For the second-order staggered velocity-stress leapfrog
  the field  v = A cos(k x - w t),  sigma = B cos(k x - w t)  (sampled at each
  field's staggered position and time level) is an exact solution when
      sin(w dt / 2) = c (dt / h) sin(k h / 2),   B = -M A dt sin(k h/2) / (h sin(w dt/2)),
  with M = lambda + 2 mu, c = cp for P waves and M = mu, c = cs for S waves.
  Nine cases (3 directions x 3 polarisations) exercise all 18 derivative terms,
  every staggering offset, both moduli, 1/rho, dt and the per-axis spacing.
* 1-D reductions in media varying along one axis, against an independent 1-D
  staggered scheme: pins which neighbours the material averages use.
* C-PML: traces near the layer match an enlarged domain, and far better than
  with rigid walls; energy decays once the wave has left.
* Exact invariances: zero in -> zero out, power-of-two linearity (bit-exact),
  translation invariance before the wave reaches the layers (bit-exact).
* Energy: closed forms for uniaxial stress, pure shear, hydrostatic pressure.
"""
import math

import numpy as np
import pytest

import oracle
from ela3d_amd import (COMPAT_GHOST, COMPAT_REFERENCE, SRC_EXPLOSIVE, SRC_FORCE, SRC_FORCE3,
                       SRC_REFERENCE_SIGMA, WAVELET_GAUSSIAN, WAVELET_RICKER, Config)

VP, VS, RHO = np.float32(3300.0), np.float32(3300.0 / 1.732), np.float32(2800.0)

# staggered position of each field in cells (x, y, z), read off the update stencils:
# sigma_xx(i) uses vx(i+1)-vx(i); vx uses (rho(i)+rho(i-1))/2; sigma_xy averages (i-1,j),(i,j+1) ...
POS = {0: (-0.5, 0.0, 0.0), 1: (0.0, 0.5, 0.0), 2: (0.0, 0.0, 0.5),
       3: (0.0, 0.0, 0.0), 4: (0.0, 0.0, 0.0), 5: (0.0, 0.0, 0.0),
       6: (-0.5, 0.5, 0.0), 7: (-0.5, 0.0, 0.5), 8: (0.0, 0.5, 0.5)}
SHEAR = {(0, 1): 6, (1, 0): 6, (0, 2): 7, (2, 0): 7, (1, 2): 8, (2, 1): 8}


def base_cfg(n=40, **kw):
    c = dict(nx=n, ny=n, nz=n, dx=10.0, dy=12.0, dz=8.0, dt=1.0e-3, nstep=1000, npoints_pml=0,
             factor=0.0, dirichlet=False, isrc=n // 2, jsrc=n // 2, ksrc=n // 2,
             receivers=np.zeros((0, 3), np.int32), energy_every=0)
    c.update(kw)
    return Config(**c)


def homog(c):
    shape = (c.nz, c.ny, c.nx)
    return (np.full(shape, VP, np.float32), np.full(shape, VS, np.float32),
            np.full(shape, RHO, np.float32))


def grid_coords(c):
    k, j, i = np.meshgrid(np.arange(c.nz), np.arange(c.ny), np.arange(c.nx), indexing="ij")
    return (i, j, k)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-4)])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_discrete_plane_wave(axis, comp, dtype, tol):
    c = base_cfg()
    T = 12
    h = (c.dx, c.dy, c.dz)[axis]
    rho, vp, vs = float(RHO), float(VP), float(VS)
    lam, mu = rho * (vp * vp - 2 * vs * vs), rho * vs * vs
    P = comp == axis
    M = lam + 2 * mu if P else mu
    kh = 2 * math.pi / 10.0
    k = kh / h
    s = math.sqrt(M / rho) * c.dt / h * math.sin(kh / 2)
    w = 2 * math.asin(s) / c.dt
    A = 1.0
    B = -M * A * c.dt * math.sin(kh / 2) / (h * math.sin(w * c.dt / 2))
    idx = grid_coords(c)
    vfid = comp
    sfid = 3 + axis if P else SHEAR[(axis, comp)]

    def field(fid, amp, n, is_v):
        x = (idx[axis] + POS[fid][axis]) * h
        t = n * c.dt if is_v else (n - 0.5) * c.dt
        return amp * np.cos(k * x - w * t)

    st = oracle.State(c, dtype)
    st.f[vfid][...] = field(vfid, A, 0, True)
    st.f[sfid][...] = field(sfid, B, 0, False)
    if P:
        for other in range(3):
            if other != axis:
                st.f[3 + other][...] = field(3 + other, B * lam / M, 0, False)
    oracle.run(c, homog(c), T, st=st, dtype=dtype)
    inner = tuple(slice(T + 2, n - T - 2) for n in (c.nz, c.ny, c.nx))
    exp_v = field(vfid, A, T, True)
    exp_s = field(sfid, B, T, False)
    assert np.max(np.abs(st.f[vfid] - exp_v)[inner]) <= tol * abs(A)
    assert np.max(np.abs(st.f[sfid] - exp_s)[inner]) <= tol * abs(B)
    for q in range(9):  # nothing else is excited
        if q not in (vfid, sfid) and not (P and q in (3, 4, 5)):
            assert np.max(np.abs(st.f[q][inner])) <= tol * max(abs(A), abs(B))


def _reduce_1d(c, axis, comp, model, steps, v0, shift=False):
    """Independent 1-D staggered scheme of the 3-D loop in a medium varying along `axis`,
    for the wave polarised along `comp`.  Each branch is read off the reference loop for
    the one stress and one velocity the reduction keeps (1-D index x along `axis`):
    which neighbours the modulus and the density are averaged over, which cells are
    updated, and the sign of each difference.  `mu[x]` is the average over cells x, x+1
    (both its factors averaged, as the reference does: rho_bar * cs_bar**2).  With
    `shift`, every material factor is taken one cell over (a deliberate mistake)."""
    cp, cs, rh = (np.asarray(m, np.float64) for m in model)
    sl = [c.nz // 2, c.ny // 2, c.nx // 2]
    sl[2 - axis] = slice(None)
    cp1, cs1, r1 = cp[tuple(sl)], cs[tuple(sl)], rh[tuple(sl)]
    n = len(cp1)
    h = (c.dx, c.dy, c.dz)[axis]
    dt = c.dt
    M = r1 * cp1 * cp1                              # lambda + 2 mu at the cell (CPU:385-387)
    mu = ((cs1[:-1] + cs1[1:]) / 2) ** 2 * ((r1[:-1] + r1[1:]) / 2)
    rbar = (r1[:-1] + r1[1:]) / 2                   # rbar[x]: density averaged over x, x+1
    if shift:  # the mistake the test must catch: every material factor one cell over
        M, mu, rbar, r1 = np.roll(M, 1), np.roll(mu, 1), np.roll(rbar, 1), np.roll(r1, 1)
    v = v0.astype(np.float64).copy()
    s = np.zeros(n)
    for _ in range(steps):
        if (axis, comp) == (0, 0):
            # sigma_xx(i), i in [0,n-2], from vx(i+1) - vx(i), modulus at i (CPU:384-388)
            s[:-1] += dt * M[:-1] * (v[1:] - v[:-1]) / h
            # vx(i), i in [1,n-1], from sxx(i) - sxx(i-1), rho over (i, i-1) (CPU:512-516)
            v[1:] += dt / rbar * (s[1:] - s[:-1]) / h
        elif (axis, comp) == (1, 1):
            # sigma_yy(j), j in [1,n-1], from vy(j) - vy(j-1) (CPU:389)
            s[1:] += dt * M[1:] * (v[1:] - v[:-1]) / h
            # vy(j), j in [0,n-2], from syy(j+1) - syy(j), rho over (j, j+1) (CPU:541-545)
            v[:-1] += dt / rbar * (s[1:] - s[:-1]) / h
        elif (axis, comp) == (2, 2):
            # sigma_zz(k), k in [1,n-1], from vz(k) - vz(k-1) (CPU:390)
            s[1:] += dt * M[1:] * (v[1:] - v[:-1]) / h
            # vz(k), k in [0,n-2], from szz(k+1) - szz(k), rho over (k, k+1) (CPU:571-576)
            v[:-1] += dt / rbar * (s[1:] - s[:-1]) / h
        elif (axis, comp) == (0, 1):
            # sigma_xy(i), i in [1,n-1], from vy(i) - vy(i-1); mu over (i-1, j), (i, j+1)
            # = cells i-1, i along x (CPU:423-431)
            s[1:] += dt * mu * (v[1:] - v[:-1]) / h
            # vy(i), i in [0,n-2], from sxy(i+1) - sxy(i); rho over (j, j+1) = rho(i) (CPU:541-544)
            v[:-1] += dt / r1[:-1] * (s[1:] - s[:-1]) / h
        elif (axis, comp) == (0, 2):
            # sigma_xz(i), i in [1,n-1], from vz(i) - vz(i-1); mu over (i-1, k), (i, k+1)
            # = cells i-1, i along x (CPU:453-460)
            s[1:] += dt * mu * (v[1:] - v[:-1]) / h
            # vz(i), i in [0,n-2], from sxz(i+1) - sxz(i); rho over (k, k+1) = rho(i) (CPU:571-574)
            v[:-1] += dt / r1[:-1] * (s[1:] - s[:-1]) / h
        elif (axis, comp) == (1, 0):
            # sigma_xy(j), j in [0,n-2], from vx(j+1) - vx(j); mu over (i-1, j), (i, j+1)
            # = cells j, j+1 along y (CPU:423-432)
            s[:-1] += dt * mu * (v[1:] - v[:-1]) / h
            # vx(j), j in [1,n-1], from sxy(j) - sxy(j-1); rho over (i, i-1) = rho(j) (CPU:512-515)
            v[1:] += dt / r1[1:] * (s[1:] - s[:-1]) / h
        elif (axis, comp) == (1, 2):
            # sigma_yz(j), j in [0,n-2], from vz(j+1) - vz(j); mu over (j+1, k), (j, k+1)
            # = cells j+1, j along y (CPU:480-487)
            s[:-1] += dt * mu * (v[1:] - v[:-1]) / h
            # vz(j), j in [1,n-1], from syz(j) - syz(j-1); rho over (k, k+1) = rho(j) (CPU:571-575)
            v[1:] += dt / r1[1:] * (s[1:] - s[:-1]) / h
        elif (axis, comp) == (2, 0):
            # sigma_xz(k), k in [0,n-2], from vx(k+1) - vx(k); mu over (i-1, k), (i, k+1)
            # = cells k, k+1 along z (CPU:453-461)
            s[:-1] += dt * mu * (v[1:] - v[:-1]) / h
            # vx(k), k in [1,n-1], from sxz(k) - sxz(k-1); rho over (i, i-1) = rho(k) (CPU:512-516)
            v[1:] += dt / r1[1:] * (s[1:] - s[:-1]) / h
        elif (axis, comp) == (2, 1):
            # sigma_yz(k), k in [0,n-2], from vy(k+1) - vy(k); mu over (j+1, k), (j, k+1)
            # = cells k, k+1 along z (CPU:480-488)
            s[:-1] += dt * mu * (v[1:] - v[:-1]) / h
            # vy(k), k in [1,n-1], from syz(k) - syz(k-1); rho over (j, j+1) = rho(k) (CPU:541-545)
            v[1:] += dt / r1[1:] * (s[1:] - s[:-1]) / h
        else:
            raise ValueError((axis, comp))
    return v


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_heterogeneous_1d_reduction(axis, comp):
    """Every (axis, polarisation) pair in a medium whose cp, cs and rho all vary along
    one axis: the 3-D oracle's middle line equals the independent 1-D scheme, which
    pins the neighbour pair of each material average (a shift of any average to the
    other side changes the result by far more than the 1e-12 bar)."""
    T = 30
    nline, nside = 64, 2 * T + 9
    dims = [nside, nside, nside]
    dims[axis] = nline
    c = base_cfg(n=8, nx=dims[0], ny=dims[1], nz=dims[2], isrc=dims[0] // 2, jsrc=dims[1] // 2,
                 ksrc=dims[2] // 2)
    rng = np.random.default_rng(5 + 3 * axis + comp)
    prof = lambda lo, hi: (lo + (hi - lo) * rng.random(nline)).astype(np.float32)
    vp1, r1 = prof(3000, 4200), prof(2000, 2900)
    vs1 = (vp1 / prof(1.6, 1.9)).astype(np.float32)
    shape = (c.nz, c.ny, c.nx)
    bshape = [1, 1, 1]
    bshape[2 - axis] = nline
    model = tuple(np.broadcast_to(a.reshape(bshape), shape).copy() for a in (vp1, vs1, r1))
    x = np.arange(nline)
    v0 = np.exp(-((x - nline / 2) / 4.0) ** 2)
    st = oracle.State(c, np.float64)
    st.f[comp][...] = v0.reshape(bshape)
    oracle.run(c, model, T, st=st, dtype=np.float64)
    ref = _reduce_1d(c, axis, comp, model, T, v0)
    sl = [c.nz // 2, c.ny // 2, c.nx // 2]
    sl[2 - axis] = slice(None)
    got = st.f[comp][tuple(sl)]
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    # power of the check: the same scheme with every average and cell factor taken one
    # cell over (the neighbour-shift mistake) misses by orders of magnitude more
    bad = _reduce_1d(c, axis, comp, model, T, v0, shift=True)
    assert np.max(np.abs(got - bad)) > 1e-4 * np.max(np.abs(ref))


@pytest.mark.parametrize("comp", [0, 1])
def test_ghost_top_plane_closed_form(comp):
    """With the ghost quirk (CPU:445: k up to NIMZ, so ztop reads a plane past the
    grid that is never written), sigma_xz / sigma_yz on the top plane k = nz-1 are
    updated against zero velocity, zero cs and zero rho above.  One step from a single
    excited vx (or vy) cell on the top plane, no C-PML, in a random medium:
      sigma_xz(i, j, top) = dt * ((rho(i-1)+0)/2) * ((cs(i-1)+0)/2)**2 * (0 - vx)/dz
      sigma_yz(i, j, top) = dt * ((rho(j+1)+0)/2) * ((cs(j+1)+0)/2)**2 * (0 - vy)/dz
    and without the quirk the top plane is not updated at all."""
    n = 12
    rng = np.random.default_rng(11 + comp)
    shape = (n, n, n)
    model = ((3000 + 500 * rng.random(shape)).astype(np.float32),
             (1600 + 300 * rng.random(shape)).astype(np.float32),
             (2000 + 800 * rng.random(shape)).astype(np.float32))
    cs, rho = (np.asarray(m, np.float64) for m in model[1:])
    i, j, k = 5, 6, n - 1
    amp = 0.37
    for flags, on in ((COMPAT_GHOST, True), (0, False)):
        c = base_cfg(n=n, compat_flags=flags)
        st = oracle.State(c, np.float64)
        st.f[comp][k, j, i] = amp
        oracle.stress(c, model, st)
        q = 7 if comp == 0 else 8
        got = st.f[q][k]
        exp = np.zeros((n, n))
        if on:
            if comp == 0:
                exp[j, i] = c.dt * (rho[k, j, i - 1] / 2) * (cs[k, j, i - 1] / 2) ** 2 * (-amp) / c.dz
            else:
                exp[j, i] = c.dt * (rho[k, j + 1, i] / 2) * (cs[k, j + 1, i] / 2) ** 2 * (-amp) / c.dz
        assert np.allclose(got, exp, rtol=1e-13, atol=0.0), (flags, got[j, i], exp[j, i])
        # the plane below sees the same cell as its ztop neighbour: a regular update
        below = st.f[q][k - 1]
        if comp == 0:
            mu = ((rho[k - 1, j, i - 1] + rho[k, j, i]) / 2) * ((cs[k - 1, j, i - 1] + cs[k, j, i]) / 2) ** 2
        else:
            mu = ((rho[k - 1, j + 1, i] + rho[k, j, i]) / 2) * ((cs[k - 1, j + 1, i] + cs[k, j, i]) / 2) ** 2
        assert below[j, i] == pytest.approx(c.dt * mu * amp / c.dz, rel=1e-13)


def _point_source_cfg(n, src, rec, npml, dirichlet=True):
    return Config(nx=n, ny=n, nz=n, dx=10.0, dy=10.0, dz=10.0, dt=1.6e-3, nstep=150,
                  npoints_pml=npml, f0=15.0, factor=1e7, src_kind=SRC_FORCE,
                  wavelet=WAVELET_RICKER, isrc=src, jsrc=src, ksrc=src,
                  receivers=np.asarray([rec], np.int32), dirichlet=dirichlet, energy_every=1)


def test_cpml_absorbs_like_an_unbounded_domain():
    small = _point_source_cfg(61, 30, (48, 30, 30), 10)
    big = _point_source_cfg(121, 60, (78, 60, 60), 10)
    rigid = _point_source_cfg(61, 30, (48, 30, 30), 0)
    s_small, e_small, _ = oracle.run(small, homog(small), threads=8)
    s_big, _, _ = oracle.run(big, homog(big), threads=8)
    s_rigid, _, _ = oracle.run(rigid, homog(rigid), threads=8)
    ref = s_big[:, 0, :]
    err_pml = np.linalg.norm(s_small[:, 0, :] - ref) / np.linalg.norm(ref)
    err_rigid = np.linalg.norm(s_rigid[:, 0, :] - ref) / np.linalg.norm(ref)
    assert err_pml < 0.05, err_pml
    assert err_pml < 0.1 * err_rigid, (err_pml, err_rigid)


def test_energy_conserved_without_pml():
    """Rigid walls, no layer: the cell-local energy (v and sigma half a step apart)
    oscillates by a few per cent but has no drift -- a wrong sign or modulus grows or decays."""
    c = _point_source_cfg(41, 20, (25, 20, 20), 0).replace(nstep=1050, wavelet=WAVELET_GAUSSIAN,
                                                           f0=20.0, dirichlet=True)
    _, e, _ = oracle.run(c, homog(c), threads=8)
    tot = e.sum(axis=1)
    after = tot[150:]  # the Gaussian pulse is over (t0 = 60 steps)
    assert np.max(after) / np.min(after) < 1.06
    means = after.reshape(6, 150).mean(axis=1)
    assert np.max(means) / np.min(means) < 1.01


def test_linearity_power_of_two_bit_exact():
    c = _point_source_cfg(33, 16, (20, 16, 16), 6).replace(nstep=40)
    a, ea, _ = oracle.run(c, homog(c))
    b, eb, _ = oracle.run(c.replace(factor=2e7), homog(c))
    assert np.array_equal(2.0 * a, b)
    assert np.array_equal(4.0 * ea, eb)


def test_zero_source_zero_state():
    c = _point_source_cfg(25, 12, (14, 12, 12), 4).replace(factor=0.0, nstep=10)
    _, _, st = oracle.run(c, homog(c))
    assert all(not np.any(f) for f in st.f + st.m)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_translation_invariance_before_the_layers(dtype):
    T = 18
    a = _point_source_cfg(61, 30, (33, 31, 29), 10).replace(nstep=T)
    b = a.replace(isrc=31, receivers=np.asarray([(34, 31, 29)], np.int32))
    sa, _, _ = oracle.run(a, homog(a), dtype=dtype)
    sb, _, _ = oracle.run(b, homog(b), dtype=dtype)
    assert np.any(sa)
    assert np.array_equal(sa, sb)


def test_source_injection_one_step_from_rest():
    n = 21
    rng = np.random.default_rng(3)
    shape = (n, n, n)
    model = ((3000 + 500 * rng.random(shape)).astype(np.float32),
             (1600 + 100 * rng.random(shape)).astype(np.float32),
             (2000 + 800 * rng.random(shape)).astype(np.float32))
    cp, cs, rho = (m.astype(np.float64) for m in model)
    i, j, k = 9, 11, 10
    for kind in (SRC_REFERENCE_SIGMA, SRC_FORCE, SRC_EXPLOSIVE, SRC_FORCE3):
        c = _point_source_cfg(n, 10, (i, j, k), 4).replace(src_kind=kind, isrc=i, jsrc=j, ksrc=k,
                                                            angle_force=30.0, nstep=1)
        _, _, st = oracle.run(c, model)
        s = oracle.wavelet(c, 1)[0]
        th = math.radians(30.0)
        l2m = rho[k, j, i] * cp[k, j, i] ** 2
        dt = c.dt
        exp = np.zeros(9)
        if kind == SRC_REFERENCE_SIGMA:
            exp[3] = exp[4] = math.sin(th) * s * dt * l2m
            exp[5] = math.cos(th) * s * dt * l2m
        elif kind == SRC_FORCE:
            exp[0] = math.sin(th) * s * dt / ((rho[k, j, i] + rho[k, j, i - 1]) / 2)
            exp[1] = math.cos(th) * s * dt / ((rho[k, j, i] + rho[k, j + 1, i]) / 2)
        elif kind == SRC_EXPLOSIVE:
            exp[3] = exp[4] = exp[5] = s * dt * l2m
        else:
            exp[0] = s * dt / ((rho[k, j, i] + rho[k, j, i - 1]) / 2)
            exp[1] = s * dt / ((rho[k, j, i] + rho[k, j + 1, i]) / 2)
            exp[2] = s * dt / ((rho[k, j, i] + rho[k + 1, j, i]) / 2)
        for q in range(9):
            assert st.f[q][k, j, i] == pytest.approx(exp[q], rel=1e-12, abs=1e-300), (kind, q)
            rest = st.f[q].copy()
            rest[k, j, i] = 0.0
            assert not np.any(rest), (kind, q)


def test_receivers_sample_before_the_source():
    """A force source's own cell reads 0 at step 1: the sample precedes injection."""
    c = _point_source_cfg(21, 10, (10, 10, 10), 4).replace(nstep=2)
    seis, _, st = oracle.run(c, homog(c))
    assert seis[0, 0, 1] == 0.0
    assert seis[1, 0, 1] != 0.0


def test_energy_closed_forms():
    c = base_cfg(n=12, npoints_pml=2)
    model = homog(c)
    rho, vp, vs = float(RHO), float(VP), float(VS)
    lam, mu = rho * (vp * vp - 2 * vs * vs), rho * vs * vs
    young = mu * (3 * lam + 2 * mu) / (lam + mu)
    bulk = lam + 2 * mu / 3
    ncell = (12 - 1 - 2 * 2) ** 3
    vol = c.dx * c.dy * c.dz
    st = oracle.State(c, np.float64)
    st.f[0][...] = 0.3
    st.f[2][...] = -0.4
    e = oracle.energy(c, model, st)
    assert e[0] == pytest.approx(0.5 * rho * 0.25 * ncell * vol, rel=1e-12)
    st = oracle.State(c, np.float64)
    st.f[3][...] = 1e6
    assert oracle.energy(c, model, st)[1] == pytest.approx(0.5 * 1e12 / young * ncell * vol, rel=1e-12)
    st = oracle.State(c, np.float64)
    st.f[7][...] = 2e5
    assert oracle.energy(c, model, st)[1] == pytest.approx(0.5 * 4e10 / mu * ncell * vol, rel=1e-12)
    st = oracle.State(c, np.float64)
    for q in (3, 4, 5):
        st.f[q][...] = -3e5
    assert oracle.energy(c, model, st)[1] == pytest.approx(0.5 * 9e10 / bulk * ncell * vol, rel=1e-12)


def test_reference_mode_runs_and_differs_only_through_quirks():
    """compat = reference changes tables and the top sxz/syz plane only."""
    c = _point_source_cfg(41, 20, (25, 20, 20), 8).replace(nstep=40)
    a, _, _ = oracle.run(c, homog(c))
    b, _, _ = oracle.run(c.replace(compat_flags=COMPAT_REFERENCE), homog(c))
    # before the wave reaches any layer the two modes agree exactly
    assert np.array_equal(a[:10], b[:10])
