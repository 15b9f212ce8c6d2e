"""GPU: out-of-memory handling, forced with the library's device-memory limit.

ela3d_set_device_limit caps the device bytes the library holds at once, so an
allocation past it fails with ELA3D_E_OOM exactly where a real out-of-memory
would: shot creation reports it and leaves nothing behind, and the server fails
that one shot while the others complete with the results they have alone.
"""
import numpy as np
import pytest

from ela3d_amd import Ela3dError, Server, Shot, set_device_limit, workloads

pytestmark = pytest.mark.gpu
E_OOM = -2


def small_shots():
    specs = workloads.serving_stream(nshots=2, seed=5, edges=(32,), nstep_range=(40, 60), nrec=8)
    return [s.materialize() for s in specs]


def test_create_reports_oom_and_recovers(gpu):
    cfg, model = small_shots()[0]
    try:
        set_device_limit(1 << 20)  # 1 MiB: far below one shot
        with pytest.raises(Ela3dError) as e:
            Shot(cfg, *model)
        assert e.value.code == E_OOM and "limit" in str(e.value)
    finally:
        set_device_limit(0)
    with Shot(cfg, *model) as s:  # nothing leaked: the same shot now fits
        s.run()
        assert np.any(s.seismograms())


def test_server_fails_only_the_shot_that_does_not_fit(gpu):
    shots = small_shots()
    big_spec = workloads.serving_stream(nshots=1, seed=6, edges=(160,), nstep_range=(20, 20), nrec=8)[0]
    big = big_spec.materialize()
    alone = []
    for cfg, model in shots:
        with Shot(cfg, *model) as s:
            s.run()
            alone.append(s.seismograms())
    try:
        set_device_limit(64 << 20)  # the two small shots fit, the 160^3 one does not
        with Server(max_active=3, steps_per_round=16) as srv:
            tickets = [srv.submit(shots[0][0], *shots[0][1]), srv.submit(big[0], *big[1]),
                       srv.submit(shots[1][0], *shots[1][1])]
            infos = [srv.wait(t) for t in tickets]
            seis = [srv.result(t, release=False)[0] if infos[q]["status"] == 0 else None
                    for q, t in enumerate(tickets)]
            stats = srv.stats()
    finally:
        set_device_limit(0)
    assert infos[1]["status"] == E_OOM
    assert infos[0]["status"] == 0 and infos[2]["status"] == 0
    assert stats["failed"] == 1 and stats["done"] == 3
    assert np.array_equal(seis[0], alone[0]) and np.array_equal(seis[2], alone[1])


@pytest.mark.parametrize("storage", [0, 1])  # FP32, FP16
def test_blow_up_is_flagged_without_receivers_or_energy(gpu, storage):
    """A shot beyond Courant with no receivers and no energy sums is still reported
    unstable: the tail checks the velocities and normal stresses at the source cell every step."""
    from ela3d_amd import Config, homogeneous_model
    cfg = Config(nx=24, ny=22, nz=20, dx=10.0, dy=10.0, dz=10.0, dt=4.0e-3, nstep=400,
                 npoints_pml=3, f0=18.0, isrc=12, jsrc=11, ksrc=10, energy_every=0,
                 storage=storage)
    assert cfg.nrec == 0
    model = homogeneous_model(cfg.nx, cfg.ny, cfg.nz, 3300.0, 3300.0 / 1.732, 2800.0)
    with Shot(cfg, *model) as s:
        with pytest.raises(Ela3dError) as e:
            s.run()
        assert e.value.code == -3  # ELA3D_E_UNSTABLE
    stable = cfg.replace(dt=1.0e-3, nstep=100)
    with Shot(stable, *model) as s:
        s.run()  # a stable shot is not flagged
