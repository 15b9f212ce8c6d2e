"""CPU: the reference's file formats (raw model/snapshot files, rechorv?.txt /
recsrc?.txt text traces), the MATLAB anticline port and the ela3d_run CLI's
argument handling (no GPU needed for any of it)."""
import os
import subprocess

import numpy as np
import pytest

import ela3d_amd
from ela3d_amd import _lib

CLI = os.path.join(os.path.dirname(ela3d_amd.__file__), "bin", "ela3d_run")


def test_raw_round_trip_float32_and_float64(tmp_path):
    rng = np.random.default_rng(0)
    a = (3000 + 500 * rng.random((5, 4, 3))).astype(np.float32)
    for bpv, dt in ((4, "<f4"), (8, "<f8")):
        p = str(tmp_path / f"m{bpv}.bin")
        ela3d_amd.write_raw(p, a, bpv)
        assert os.path.getsize(p) == a.size * bpv
        raw = np.fromfile(p, dtype=dt)  # x-fastest: the C-order ravel of (nz, ny, nx)
        assert np.array_equal(raw.astype(np.float32), a.ravel())
        assert np.array_equal(ela3d_amd.read_raw(p, a.shape, bpv), a)


def test_read_raw_float64_file_written_elsewhere(tmp_path):
    v = np.linspace(2000.0, 4500.0, 60)
    p = str(tmp_path / "modelcp.bin")
    v.astype("<f8").tofile(p)
    out = ela3d_amd.read_raw(p, (3, 4, 5), 8)
    assert np.array_equal(out.ravel(), v.astype(np.float32))


def test_read_raw_errors(tmp_path):
    p = str(tmp_path / "short.bin")
    np.zeros(10, "<f4").tofile(p)
    with pytest.raises(ela3d_amd.Ela3dError) as e:
        ela3d_amd.read_raw(p, (11,), 4)
    assert e.value.code == -11
    with pytest.raises(ela3d_amd.Ela3dError) as e:
        ela3d_amd.read_raw(str(tmp_path / "missing.bin"), (1,), 4)
    assert e.value.code == -11
    with pytest.raises(ela3d_amd.Ela3dError):
        ela3d_amd.read_raw(p, (1,), 2)


def test_trace_text_matches_float_stream_output(tmp_path):
    """Values print as std::cout << float does (6 significant digits), each followed
    by a space, one line per row; these literals are that stream's output."""
    vals = np.array([[1e7, 0.1, -3.14159274, 123456789.0, 0.0, -2.5e-12]], dtype=np.float64)
    p = str(tmp_path / "t.txt")
    ela3d_amd.write_traces(p, vals)
    assert open(p).read() == "1e+07 0.1 -3.14159 1.23457e+08 0 -2.5e-12 \n"


def test_rechor_layout_and_append(tmp_path):
    rng = np.random.default_rng(3)
    seis = rng.standard_normal((4, 5, 3)) * 1e-3
    ela3d_amd.write_rechor(str(tmp_path), seis, append=True)
    ela3d_amd.write_rechor(str(tmp_path), seis[:1], append=True)  # the reference appends
    for comp, name in enumerate(("rechorvx.txt", "rechorvy.txt", "rechorvz.txt")):
        lines = open(tmp_path / name).read().split("\n")
        assert lines[-1] == "" and len(lines) == 6
        for r, line in enumerate(lines[:5]):
            assert line.endswith(" ")
            got = np.array([float(x) for x in line.split()])
            want = seis[r % 4, :, comp].astype(np.float32).astype(np.float64)
            assert got.shape == (5,)
            assert np.allclose(got, want, rtol=5e-6, atol=0)


def test_recsrc_single_line_without_newline(tmp_path):
    tr = np.zeros((4, 3))
    tr[1:, 2] = [1.5, -2.0, 3.25]
    p = str(tmp_path / "recsrcz.txt")
    ela3d_amd.write_traces(p, tr.reshape(1, -1), stride=3, offset=2, final_newline=False)
    assert open(p).read() == "0 1.5 -2 3.25 "


def test_anticline_columns():
    """Closed forms of generate_3d_model_anticline.m at n = 201: at the corner column
    Z = 0 (interfaces at k = 121, 161, 201); at the crest (x = y = pi/2) Z = 1
    (interfaces at k = 21, 61, 101, oil sandstone for 61 < k < 80)."""
    cp, cs, rh = ela3d_amd.anticline_model(201)
    k = np.arange(1, 202)
    corner = np.where(k <= 121, 1, np.where(k <= 161, 2, 3))
    crest = np.select([k <= 21, k <= 61, k < 80, k <= 101], [1, 2, 4, 3], 5)
    media = {1: (3000, 1700, 2800), 2: (3600, 1400, 2300), 3: (3000, 1700, 2800),
             4: (3600, 1700, 2800), 5: (3650, 1500, 2450)}
    for col, (i, j) in ((corner, (0, 0)), (crest, (100, 100))):
        for f, arr in enumerate((cp, cs, rh)):
            assert np.array_equal(arr[:, j, i], [media[m][f] for m in col])
    assert np.array_equal(cp, np.transpose(cp, (0, 2, 1)))  # symmetric in x and y
    assert set(np.unique(rh)) == {2300, 2450, 2800}


def run_cli(*args):
    return subprocess.run([CLI, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=60)


@pytest.mark.skipif(not os.path.exists(CLI), reason="ela3d_run not built")
def test_cli_arguments():
    r = run_cli("--help")
    assert r.returncode == 0 and "rec=grid" in r.stdout
    assert run_cli("nx").returncode == 2
    assert run_cli("nx=abc").returncode == 2
    assert run_cli("src=bogus").returncode == 2
    r = run_cli("nx=20", "ny=20", "nz=20", "npml=3", "rec=grid:40,2,5")
    assert r.returncode == 2 and "more receivers" in r.stderr
    r = run_cli("nx=20", "ny=20", "nz=20", "npml=3", "isrc=0")
    assert r.returncode == 1 and "config" in r.stderr


@pytest.mark.skipif(not os.path.exists(CLI), reason="ela3d_run not built")
def test_cli_without_gpu_fails_cleanly():
    if ela3d_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    r = run_cli("nx=20", "ny=20", "nz=20", "npml=3", "rec=grid:4,2,5", "nstep=10")
    assert r.returncode == 1 and "no GPU" in r.stderr
