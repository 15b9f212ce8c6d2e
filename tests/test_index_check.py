"""CPU: every device address of every launch shape, enumerated by a C++ program
(tests/index_check.cpp) that uses the kernels' own geometry functions (geom.hpp)
and the library's own plan and launch geometry, checked against the array sizes
a shot allocates.  Cross-checks the Python model in test_index_bounds.py on the
same shapes, and covers grids too large for it (classic, the largest grid)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
LIB = os.path.join(ROOT, "ela3d_amd", "libela3d.so")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC) or not os.path.exists(LIB):
        pytest.skip("needs hipcc and the built library")
    exe = str(tmp_path_factory.mktemp("ic") / "index_check")
    cmd = [HIPCC, "-std=c++17", "-O2", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           os.path.join(ROOT, "tests", "index_check.cpp"), f"-I{os.path.join(ROOT, 'ela3d_amd', 'csrc')}",
           f"-I{os.path.join(ROOT, 'include')}", f"-L{os.path.dirname(LIB)}", "-lela3d",
           f"-Wl,-rpath,{os.path.dirname(LIB)}", "-o", exe]
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return exe


def run(exe, grid, storage=0, compat=0, block=256, zseg=4, vec=1, block_slab=128, zseg_slab=2,
        vec_slab=1, merged=0, fused=0, tile=0, mutation=0, fs=0, strips=0):
    nx, ny, nz, npml = grid
    args = [nx, ny, nz, npml, storage, compat, block, zseg, vec, block_slab, zseg_slab, vec_slab, merged,
            fused, tile, mutation, fs, strips]
    out = subprocess.run([exe, *map(str, args)], check=True, stdout=subprocess.PIPE, text=True).stdout
    last = out.strip().splitlines()[-1]
    if last.startswith("refused"):
        return None
    f = last.split()
    assert f[0] == "accesses" and f[2] == "violations", out
    return int(f[1]), int(f[3])


SMALL = [(29, 23, 26, 4), (17, 19, 18, 3), (40, 13, 15, 4), (11, 12, 13, 0), (64, 64, 9, 3)]


@pytest.mark.parametrize("grid", SMALL)
@pytest.mark.parametrize("fs", [0, 1])
def test_split_launch_shapes(checker, grid, fs):
    for storage in (0, 1, 2):
        for compat in (0, 7):
            for block, zseg, vec in ((64, 2, 1), (256, 4, 2), (128, 32, 4), (64, 1, 4)):
                for block_slab, zseg_slab, vec_slab in ((64, 1, 1), (256, 8, 2)):
                    for merged in (0, 1):
                        r = run(checker, grid, storage, compat, block, zseg, vec, block_slab, zseg_slab,
                                vec_slab, merged, fs=fs)
                        assert r is not None and r[0] > 0 and r[1] == 0, (storage, compat, block, zseg, vec, r)


@pytest.mark.parametrize("grid", SMALL)
@pytest.mark.parametrize("fs", [0, 1])
def test_fused_launch_shapes(checker, grid, fs):
    for storage in (0, 2):
        for tile, zseg, strips in ((256, 2, 0), (512, 4, 2), (1024, 8, 0), (256, 32, 3), (768, 1, 0),
                                   (256, 3, 64)):
            r = run(checker, grid, storage, block=128, zseg=zseg, fused=1, tile=tile, fs=fs, strips=strips)
            assert r is None or r[1] == 0, (storage, tile, zseg, strips, r)


def test_workload_grids(checker):
    classic, layered, cube = (101, 641, 640, 10), (401, 401, 301, 10), (101, 101, 101, 10)
    for grid in (classic, layered, cube):
        assert run(checker, grid, block=128, zseg=2)[1] == 0
        assert run(checker, grid, block=256, zseg=16, vec=2, merged=1)[1] == 0
    for tile, zseg in ((1024, 8), (1024, 16), (768, 32), (512, 4)):
        assert run(checker, classic, block=128, zseg=zseg, fused=1, tile=tile)[1] == 0
    for strips in (1, 2, 3, 4, 6):  # tile orders of the fused step
        assert run(checker, classic, block=128, zseg=16, fused=1, tile=768, strips=strips)[1] == 0
    for zseg in (12, 24):  # the tuner's other fused segment lengths, in the default order
        assert run(checker, classic, block=128, zseg=zseg, fused=1, tile=768)[1] == 0
    assert run(checker, layered, block=128, zseg=16, fused=1, tile=1024)[1] == 0


def test_largest_grids(checker):
    """The largest grids a shot accepts (padded arrays just below 4 GiB)."""
    for grid, storage in (((1024, 1024, 1023, 10), 0), ((1024, 1024, 511, 10), 2)):
        for vec in (1, 4):
            r = run(checker, grid, storage, block=256, zseg=8, vec=vec, block_slab=256, zseg_slab=4, vec_slab=2)
            assert r[1] == 0


def test_checker_catches_wrapped_and_short_arrays(checker):
    """Offsets formed from plane 0 wrap below zero at the first row; arrays one
    plane short are overrun at the top; a fused tile one own cell short leaves a
    cell no tile owns.  All must be reported."""
    grid = (29, 23, 26, 4)
    assert run(checker, grid, vec=2, mutation=1)[1] > 0
    assert run(checker, grid, vec=2, mutation=2)[1] > 0
    assert run(checker, grid, block=128, zseg=4, fused=1, tile=256, mutation=1)[1] > 0
    assert run(checker, grid, block=128, zseg=4, fused=1, tile=256, mutation=2)[1] > 0
    assert run(checker, grid, block=128, zseg=4, fused=1, tile=256, mutation=3)[1] > 0
