"""GPU: small shots stepped together in the server's batch launches.

FP32 shots without energy sums and at most batch_cells cells step as one batch:
per step one stress launch and one velocity launch walk every member's boxes
(from a work table in device memory) and one tail launch runs a workgroup per
member.  Each member's cells do the same arithmetic as when the shot runs alone
with one-element lanes, so its seismograms are bit-identical; larger shots and
shots with energy sums run on their own streams beside the batch.
"""
import numpy as np
import pytest

from ela3d_amd import Server, Shot, set_tuning, workloads

pytestmark = pytest.mark.gpu


def specs(n, seed, edges):
    return workloads.serving_stream(nshots=n, seed=seed, edges=edges, nstep_range=(50, 140),
                                    f0_range=(25.0, 40.0), nrec=12)


def alone(cfg, model):
    set_tuning(cfg, block=128, zseg_interior=2, vec=1, block_slab=128, zseg_slab=2, vec_slab=1)
    with Shot(cfg, *model) as s:
        s.run()
        return s.seismograms()


def serve(shots, **opts):
    with Server(**opts) as srv:
        tickets = [srv.submit(cfg, *model) for cfg, model in shots]
        out = [srv.result(t) for t in tickets]
        stats = srv.stats()
    return out, stats


def test_batched_shots_match_alone(gpu):
    shots = [s.materialize() for s in specs(7, 21, (32, 40, 48))]
    shots.append([s.materialize() for s in specs(1, 22, (72,))][0])  # above batch_cells: own stream
    ref = [alone(c, m) for c, m in shots]
    out, stats = serve(shots, max_active=8, steps_per_round=24, batch_cells=56 ** 3)
    print("server stats", stats)
    assert stats["done"] == len(shots) and stats["failed"] == 0
    assert stats["batched"] == 7 and stats["batch_rounds"] > 0  # the 72^3 shot ran on its own stream
    for (seis, en, info), r, (cfg, _) in zip(out, ref, shots):
        assert info["status"] == 0
        assert np.any(r)
        assert np.array_equal(seis, r), f"{cfg.nx}^3 shot differs when batched"


def test_batch_admits_arrivals_while_running(gpu):
    """More small shots than active slots: members leave as they finish and queued
    shots join later rounds; a blown-up member fails alone."""
    sp = specs(10, 23, (32, 40))
    shots = [s.materialize() for s in sp]
    bad_cfg, bad_model = shots[3]
    shots[3] = (bad_cfg.replace(dt=bad_cfg.dt * 4.0), bad_model)  # beyond Courant: unstable
    ref = [alone(c, m) for q, (c, m) in enumerate(shots) if q != 3]
    out, stats = serve(shots, max_active=4, steps_per_round=16)
    assert stats["done"] == len(shots) and stats["failed"] == 1 and stats["batched"] == len(shots)
    assert out[3][2]["status"] == -3  # ELA3D_E_UNSTABLE
    good = [o for q, o in enumerate(out) if q != 3]
    for (seis, en, info), r in zip(good, ref):
        assert info["status"] == 0 and np.array_equal(seis, r)


def test_batched_free_surface_shots_match_alone(gpu):
    """Members with a free surface (plane 0 a box of the absorbing-layer kind, which
    the batch launch walks like any other) step bit-identically to running alone."""
    shots = [s.materialize() for s in specs(5, 31, (32, 40))]
    shots = [(c.replace(free_surface=(q % 2 == 0), ksrc=3), m) for q, (c, m) in enumerate(shots)]
    ref = [alone(c, m) for c, m in shots]
    out, stats = serve(shots, max_active=8, steps_per_round=20, batch_cells=48 ** 3)
    assert stats["done"] == len(shots) and stats["failed"] == 0 and stats["batched"] == len(shots)
    for (seis, en, info), r in zip(out, ref):
        assert info["status"] == 0 and np.any(r)
        assert np.array_equal(seis, r)
