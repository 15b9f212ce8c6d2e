"""CPU: host side of the shot server -- options, error paths without a GPU, the
seeded serving stream and the FIFO latency model of one-at-a-time execution."""
import ctypes as C

import numpy as np
import pytest

from ela3d_amd import _lib, workloads


def test_server_opts_default():
    o = _lib.ServerOpts()
    _lib.lib().ela3d_server_opts_default(C.byref(o))
    assert o.struct_size == C.sizeof(_lib.ServerOpts)
    assert o.hbm_budget_bytes == 200 * 10**9
    assert o.max_active >= 1 and o.steps_per_round >= 1 and o.use_graph == 1 and o.plain_loop == 0


@pytest.mark.parametrize("field,value", [("max_active", 0), ("steps_per_round", 0),
                                         ("hbm_budget_bytes", 0), ("struct_size", 8)])
def test_server_rejects_bad_options(field, value):
    L = _lib.lib()
    o = _lib.ServerOpts()
    L.ela3d_server_opts_default(C.byref(o))
    setattr(o, field, value)
    h = C.c_void_p()
    assert L.ela3d_server_create(C.byref(o), C.byref(h)) == -1  # ELA3D_E_INVALID
    assert not h.value


def test_server_needs_a_device():
    import ela3d_amd
    if ela3d_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    L = _lib.lib()
    o = _lib.ServerOpts()
    L.ela3d_server_opts_default(C.byref(o))
    h = C.c_void_p()
    assert L.ela3d_server_create(C.byref(o), C.byref(h)) == -5  # ELA3D_E_NODEVICE


def test_null_server_arguments():
    L = _lib.lib()
    st = C.c_int32()
    assert L.ela3d_poll(None, 1, C.byref(st)) == -1
    out = (C.c_int64 * 8)()
    assert L.ela3d_server_stats(None, out) == -1
    L.ela3d_server_destroy(None)  # no-op


def test_serving_stream_is_seeded_and_in_range():
    a = workloads.serving_stream()
    b = workloads.serving_stream()
    assert len(a) == 256
    assert [s.arrival for s in a] == [s.arrival for s in b]
    assert all(np.array_equal(x.receivers, y.receivers) for x, y in zip(a, b))
    arr = np.asarray([s.arrival for s in a])
    assert np.all(np.diff(arr) > 0)
    # Poisson at 4/s: 256 arrivals take 64 s on average (sd 4 s)
    assert 48.0 < arr[-1] < 80.0
    assert {s.edge for s in a} == set(workloads.SERVING_EDGES)
    assert all(500 <= s.nstep <= 2000 and 5.0 <= s.f0 <= 15.0 for s in a)
    assert 0.3 < np.mean([s.layered for s in a]) < 0.7
    for s in a:
        lo, hi = 12, s.edge - 12
        assert s.receivers.shape == (64, 3)
        assert s.receivers.min() >= lo and s.receivers.max() < hi
        assert all(lo <= v < hi for v in s.src)
    assert workloads.serving_stream(seed=8)[0].arrival != a[0].arrival


def test_serving_shot_configs_validate():
    for s in workloads.serving_stream(nshots=6, seed=2, edges=(64, 96)):
        cfg, model = s.materialize()
        cfg.validate()
        assert model[0].shape == (s.edge,) * 3
        p = cfg.plan()
        assert p["cells"] == s.edge ** 3 and p["nsamples"] == s.nstep
        vmax = float(model[0].max())
        courant = vmax * cfg.dt * np.sqrt(3.0) / 10.0
        assert courant == pytest.approx(0.5, rel=1e-6)


def test_fifo_latency_model():
    # idle server: latency = service time
    lat = workloads.fifo_latencies([0.0, 10.0, 20.0], [1.0, 2.0, 3.0])
    assert np.allclose(lat, [1.0, 2.0, 3.0])
    # burst: each request waits for all earlier ones
    lat = workloads.fifo_latencies([0.0, 0.0, 0.0], [1.0, 2.0, 3.0])
    assert np.allclose(lat, [1.0, 3.0, 6.0])
    # partial overlap
    lat = workloads.fifo_latencies([0.0, 0.5, 4.0], [2.0, 2.0, 1.0])
    assert np.allclose(lat, [2.0, 3.5, 1.0])
