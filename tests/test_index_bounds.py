"""CPU model of every device address the stencil kernels form, checked against
the allocation sizes (regression test for the index arithmetic of kernels.hip).

The kernels use unsigned 32-bit element offsets from the start of each field or
model allocation: o = g0 + k*nxyp + p (g0 guard elements, nxyp the padded plane
stride), neighbours at o +- 1, o +- nx, o + nxyp and up to VEC-1 beyond.  An
offset formed below zero wraps to ~2^32 and addresses far outside the buffer,
so this model reproduces the wrap exactly.  Memory variables use compact per-axis
offsets and are touched only for elements inside the workgroup's box.
"""
import numpy as np
import pytest

from ela3d_amd import workloads
from helpers import small_cfg

M32 = 1 << 32


def plane_stride(nx, ny):
    return (nx * ny + 63) // 64 * 64


def guard(nx):
    return (nx + 64 + 63) // 64 * 64


def violations(cfg, block, vec, zi, zs, base_at_guard=False, block_slab=None, vec_slab=None):
    """Out-of-bounds accesses of the interior launch (block, vec, zseg zi) and the
    absorbing-layer launch (block_slab, vec_slab; zseg zs on the z layers, zi on
    the y layers) over one step."""
    block_slab = block if block_slab is None else block_slab
    vec_slab = min(vec, 2) if vec_slab is None else vec_slab
    p = cfg.plan()
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    nxy, nxyp, G = nx * ny, plane_stride(nx, ny), guard(nx)
    alloc = nz * nxyp + 2 * G
    lo, hi = p["region_lo"], p["region_hi"]
    w = [lo[a] + ((nx, ny, nz)[a] - hi[a]) for a in range(3)]
    memsize = [nz * ny * w[0], nz * w[1] * nx, w[2] * ny * nx]
    pa, pb = (lo[1] * nx + 63) // 64 * 64, (hi[1] * nx) // 64 * 64
    if pb <= pa:
        pa = pb = nxy
    zl = min(1, hi[2]) if (cfg.free_surface and lo[2] < 1) else lo[2]  # a free surface is a box of its own
    boxes = []
    for (p0, p1, k0, k1, z, mode) in ((0, nxy, 0, zl, zs, 1), (0, nxy, hi[2], nz, zs, 1),
                                      (0, pa, zl, hi[2], zi, 1), (pb, nxy, zl, hi[2], zi, 1),
                                      (pa, pb, zl, hi[2], zi, 0)):
        if p1 > p0 and k1 > k0:
            boxes.append((p0, p1, k0, k1, z, mode))
    vec_i = vec
    bad = []
    # base_at_guard models pointers placed at plane 0 (the addressing that wrapped)
    cell = (lambda k, q: (k * nxyp + q) % M32) if base_at_guard else (lambda k, q: (G + k * nxyp + q) % M32)
    shift = G if base_at_guard else 0

    def fld(o, tag, vec):
        for e in range(vec):
            if not (shift + (o + e) % M32 < alloc):
                bad.append(tag)

    for (p0b, p1b, k0, k1, z, mode) in boxes:
        blk, vec = (block, vec_i) if mode == 0 else (block_slab, vec_slab)
        per = blk * vec
        for chunk in range((p1b - p0b + per - 1) // per):
            for zsg in range((k1 - k0 + z - 1) // z):
                kb = k0 + zsg * z
                ke = min(kb + z, k1)
                for t in range(blk):
                    p0 = p0b + (chunk * blk + t) * vec
                    if p0 >= p1b:
                        continue
                    assert p0 % vec == 0
                    els = []
                    for e in range(vec):
                        j, i = divmod(p0 + e, nx)
                        act = p0 + e < p1b
                        cx = -1 if not act else (i if i < lo[0] else (lo[0] + i - hi[0] if i >= hi[0] else -1))
                        cy = -1
                        if mode == 1 and act:
                            cy = j if j < lo[1] else (lo[1] + j - hi[1] if j >= hi[1] else -1)
                        els.append((i, j, act, cx, cy))
                    fld(cell(kb, p0), "queue", vec)
                    if kb >= 1:
                        fld((cell(kb, p0) - nxyp) % M32, "queue-1", vec)
                    for k in range(kb, ke):
                        o = cell(k, p0)
                        dU = 0 if k == nz - 1 else nxyp
                        for d in (0, dU, 1, -1, nx, -nx):
                            fld((o + d) % M32, ("neighbour", k, p0, d), vec)
                        cz = -1
                        if mode == 1:
                            cz = k if k < lo[2] else (lo[2] + k - hi[2] if k >= hi[2] else -1)
                        for (i, j, act, cx, cy) in els:
                            if cx >= 0 and ((k * ny + j) * w[0] + cx) % M32 >= memsize[0]:
                                bad.append(("mx", k, i, j))
                            if cy >= 0 and ((k * w[1] + cy) * nx + i) % M32 >= memsize[1]:
                                bad.append(("my", k, i, j))
                            if cz >= 0 and act and ((cz * ny + j) * nx + i) % M32 >= memsize[2]:
                                bad.append(("mz", k, i, j))
    return bad


CASES = [(29, 23, 26, 4), (17, 19, 18, 3), (40, 13, 15, 4), (11, 12, 13, 0), (64, 64, 9, 3)]


@pytest.mark.parametrize("shape", CASES)
@pytest.mark.parametrize("vec", [1, 2, 4])
def test_every_access_in_bounds(shape, vec):
    nx, ny, nz, npml = shape
    for compat in (0, 7):
        cfg = small_cfg(nx=nx, ny=ny, nz=nz, npml=npml, compat=compat)
        for block, zi, zs in ((64, 2, 2), (256, 4, 4), (128, 32, 8)):
            # split launches, and the merged launch (absorbing layers with the interior shape)
            for vs, bs in ((1, 64), (2, 256), (2, 128), (vec, block)):
                assert violations(cfg, block, vec, zi, zs, block_slab=bs, vec_slab=vs) == []


def test_cube_shape_in_bounds():
    cfg, _ = workloads.homogeneous_cube(n=45, nstep=1)
    for vec in (1, 2, 4):
        assert violations(cfg, 256, vec, 4, 4) == []


def test_model_catches_wrapped_offsets():
    """Indexing from plane 0 with unsigned offsets wraps at the first row (o - 1, o - nx)."""
    bad = violations(small_cfg(), 256, 2, 4, 4, base_at_guard=True)
    assert ("neighbour", 0, 0, -1) in bad and ("neighbour", 0, 0, -29) in bad


def fused_violations(cfg, tile, zseg, elem=4, ldsreg=None):
    """Every global and LDS index the fused interior kernel forms (k_fused and its
    helpers in fused.hip: one thread per region cell, own cells first, then the
    halo row below and the row above; fused_cell in geom.hpp), for all tiles and z
    segments, checked against the allocation, memory-variable and LDS sizes.  Also
    checks that the region threads cover the region exactly once and that every
    LDS cell an own cell reads was written in that plane."""
    p = cfg.plan()
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    nxy, nxyp, G = nx * ny, plane_stride(nx, ny), guard(nx)
    alloc = nz * nxyp + 2 * G
    lo, hi = p["region_lo"], p["region_hi"]
    w0 = lo[0] + nx - hi[0]
    memsize0 = nz * ny * w0
    pa, pb = (lo[1] * nx + 63) // 64 * 64, (hi[1] * nx) // 64 * 64
    zlo, zhi = lo[2], hi[2]
    if cfg.free_surface and zlo < 1:  # plane 0 is a box of the absorbing-layer launches
        zlo = 1
    if pb <= pa or zhi <= zlo or lo[1] < 1 or zlo < 1 or hi[1] > ny - 1 or hi[2] > nz - 1:
        return None  # no interior box, or a face without absorbing layer: refused (fused_unsupported)
    maxt = 512 if elem == 8 else 1024
    own_full = (tile - 2 * nx) // 32 * 32  # geom.hpp fused_own: a multiple of 32
    if tile < 64 or tile > maxt or tile % 64 or own_full < 64:
        return None  # the library refuses this tile (fused_lds_bytes)
    LR = tile if ldsreg is None else ldsreg
    bad = []
    nchunk = (pb - pa + own_full - 1) // own_full
    nzs = (zhi - zlo + zseg - 1) // zseg
    t = np.arange(tile)

    def cell(k, q):
        return G + k * nxyp + q

    def fld(o, tag):
        o = np.asarray(o)
        if np.any((o < 0) | (o >= alloc)):
            bad.append(tag)

    def mem(k, q):  # x memory variables: offset 0 outside the layer
        j, i = np.divmod(q, nx)
        cx = np.where(i < lo[0], i, np.where(i >= hi[0], lo[0] + i - hi[0], -1))
        if np.any(j >= ny):
            bad.append(("row", k))
        if w0 > 0:
            om = np.where(cx >= 0, (k * ny + j) * w0 + cx, 0)
            if np.any((om < 0) | (om >= memsize0)):
                bad.append(("mem", k))

    def load(q, k):  # fload
        if not (0 <= k < nz - 1):
            bad.append(("plane", k))
        o = cell(k, q)
        for d in (0, nxyp, 1, -1, nx, -nx):
            fld(o + d, ("stress", k, d))
        mem(k, q)

    def qinit(q, k):
        if not (1 <= k < nz):
            bad.append(("qplane", k))
        fld(cell(k, q), ("queue", k))
        fld(cell(k, q) - nxyp, ("queue-1", k))

    for chunk in range(nchunk):
        q0 = pa + chunk * own_full
        n = min(own_full, pb - q0)
        c = np.where(t < n, t, np.where(t < n + nx, t - n - nx, np.where(t < n + 2 * nx, t - nx, -10 ** 9)))
        act = c != -10 ** 9
        own = t < n
        pc = q0 + np.where(act, c, 0)
        if np.any((pc < 0) | (pc >= nxy)):
            bad.append(("cell", chunk))
        if not np.array_equal(np.sort(c[act]), np.arange(-nx, n + nx)):
            bad.append(("coverage", chunk))
        if np.any((c[act] + nx < 0) | (c[act] + nx >= LR)):
            bad.append(("lds-write", chunk))
        inb = act & (pc >= pa) & (pc < pb)
        for zs in range(nzs):
            kb = zlo + zs * zseg
            ke = min(kb + zseg, zhi)

            def computes(m):
                if m < kb - 1 or m > ke or m < zlo or m >= zhi:
                    return np.zeros(tile, bool)
                return own.copy() if m in (kb - 1, ke) else inb.copy()

            m0 = kb - 1 if computes(kb - 1).any() else kb
            first = computes(kb - 1) if m0 == kb - 1 else computes(kb)
            qinit(pc[first], m0)
            if m0 == kb - 1:  # the queue loads it does not need there come from plane kb
                qinit(pc[first], kb)
            # own cells whose plane kb-1 is not computed start their queue at kb
            qinit(pc[own & ~computes(kb - 1)], kb)
            for m in range(kb - 1, ke + 1):
                cm = computes(m)
                # loads a boundary plane does not need move one plane inwards (dA, dB)
                shift = 1 if m == kb - 1 else (-1 if m == ke else 0)
                for s in {0, shift}:
                    load(pc[cm], m + s)
                    load(pc[act & ~cm], kb + s)  # the safe offset of the others
                region = kb <= m < ke
                fresh = ~cm & (own | (act if region else np.zeros(tile, bool)))
                fld(cell(m, pc[fresh]), ("fresh", m))
                k = m - 1
                if m > kb:
                    ownp = pc[own]
                    mem(k, ownp)
                    co = c[own]
                    for d in (0, -1, 1, -nx, nx):
                        idx = co + d + nx
                        if np.any((idx < 0) | (idx >= LR)):
                            bad.append(("lds", k, d))
                        if not np.all(np.isin(co + d, c[act])):
                            bad.append(("lds-unwritten", k, d))
                    fld(cell(k, ownp), ("velocity", k))
    return bad


@pytest.mark.parametrize("shape", [(29, 23, 26, 4), (40, 13, 15, 4), (17, 19, 18, 3), (64, 64, 9, 3)])
@pytest.mark.parametrize("tile,zseg", [(256, 2), (512, 4), (1024, 8), (256, 32)])
def test_fused_accesses_in_bounds(shape, tile, zseg):
    nx, ny, nz, npml = shape
    for compat in (0, 7):
        cfg = small_cfg(nx=nx, ny=ny, nz=nz, npml=npml, compat=compat)
        assert fused_violations(cfg, tile, zseg) in ([], None)
        assert fused_violations(cfg, tile, zseg, elem=8) in ([], None)
    cfg = small_cfg(nx=nx, ny=ny, nz=nz, npml=npml)
    if nx <= 96:
        assert fused_violations(cfg, tile, zseg) == [] or tile - 2 * nx < 64


def test_fused_cube_and_classic_in_bounds():
    cfg, _ = workloads.homogeneous_cube(n=64, nstep=1)
    for tile, zseg in ((512, 8), (1024, 16)):
        assert fused_violations(cfg, tile, zseg) == []
    cfg, _ = workloads.classic(nstep=1)
    assert fused_violations(cfg, 1024, 16) == []
    assert fused_violations(cfg, 768, 32) == []


def test_fused_layered_refused_or_in_bounds():
    """nx = 401 leaves 222 own cells in a 1024-thread tile; 512 threads leave none."""
    cfg, _ = workloads.layered(nstep=1)
    assert fused_violations(cfg, 1024, 16) == []
    assert fused_violations(cfg, 512, 16) is None


def test_fused_model_catches_short_lds_plane():
    cfg = small_cfg(nx=29, ny=23, nz=26, npml=4)
    bad = fused_violations(cfg, 256, 4, ldsreg=256 - 29)
    assert any(b[0] == "lds-write" for b in bad) and any(b[0] == "lds" for b in bad)


def xcd_block(b, n):
    """The renumbering of k_stress/k_velocity/k_fused (kernels.hip): XCD x owns a
    contiguous run of logical blocks."""
    q, r = divmod(n, 8)
    x, l = b % 8, b // 8
    return x * q + min(x, r) + l


def tile_of(nchunk, nzs, local):
    """The strip-major tile order of kernels.hip: up to 8 strips of at least 8 chunks."""
    ns = min(8, max(1, nchunk // 8))
    w, r = divmod(nchunk, ns)
    cp = local // nzs
    s = cp // (w + 1) if cp < r * (w + 1) else r + (cp - r * (w + 1)) // w
    ws = w + (1 if s < r else 0)
    c0 = s * w + min(s, r)
    off = local - c0 * nzs
    return c0 + off % ws, off // ws


@pytest.mark.parametrize("nzs", [1, 2, 3, 5, 310])
def test_block_order_is_a_bijection(nzs):
    """Every (chunk, z segment) tile of a box is run by exactly one workgroup, for
    narrow boxes (one strip), wide ones (8 strips) and everything between."""
    for nchunk in list(range(1, 70)) + [490, 506]:
        n = nchunk * nzs
        tiles = {tile_of(nchunk, nzs, xcd_block(b, n)) for b in range(n)}
        assert tiles == {(c, z) for c in range(nchunk) for z in range(nzs)}
