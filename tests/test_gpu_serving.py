"""GPU: the continuous-batching shot server.

Every shot's seismograms and energy are required to be bit-identical whether the
shot runs alone (``Shot.run``) or among others in the server, under any
admission order, budget or round length: the kernels are deterministic (the
energy reduction sums per-block partials in a fixed order) and shots share no
device memory.  One served shot is also checked against the FP64 oracle at the
1e-4 seismogram bar.
"""
import numpy as np
import pytest

import oracle
from ela3d_amd import Ela3dError, Server, Shot, workloads
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def mixed_specs(n=8, seed=3):
    """Small mixed-size shots of the serving stream's kind (homogeneous and layered)."""
    return workloads.serving_stream(nshots=n, seed=seed, edges=(32, 40, 48, 56),
                                    nstep_range=(60, 160), f0_range=(25.0, 40.0), nrec=16)


def materialize(spec, **kw):
    cfg, model = spec.materialize()
    return cfg.replace(energy_every=10, **kw), model


def alone(cfg, model):
    with Shot(cfg, *model) as s:
        s.run()
        return s.seismograms(), s.energy()


def serve(shots, **opts):
    with Server(**opts) as srv:
        tickets = [srv.submit(cfg, *model) for cfg, model in shots]
        out = [srv.result(t) for t in tickets]
        stats = srv.stats()
    return out, stats


def test_mix_matches_alone(gpu):
    shots = [materialize(s) for s in mixed_specs()]
    ref = [alone(c, m) for c, m in shots]
    out, stats = serve(shots, max_active=4, steps_per_round=16)
    print("server stats", stats)
    assert stats["done"] == len(shots) and stats["failed"] == 0
    for (seis, en, info), (rs, re), (cfg, _) in zip(out, ref, shots):
        assert info["status"] == 0
        assert seis.shape == rs.shape and en.shape == re.shape
        assert np.array_equal(seis, rs), f"shot {cfg.nx}^3 differs in the mix"
        assert np.array_equal(en, re)
        assert np.abs(rs).max() > 0


def test_admission_under_budget(gpu):
    shots = [materialize(s) for s in mixed_specs(6, seed=11)]
    need = max(c.plan()["device_bytes"] for c, _ in shots)
    budget = int(2.5 * need)
    out, stats = serve(shots, max_active=16, steps_per_round=8, hbm_budget_bytes=budget)
    print(f"budget {budget} B, peak {stats['peak_bytes']} B")
    assert stats["peak_bytes"] <= budget
    assert stats["done"] == len(shots) and stats["failed"] == 0
    for (seis, en, _), (c, m) in zip(out, shots):
        rs, re = alone(c, m)
        assert np.array_equal(seis, rs) and np.array_equal(en, re)


def test_plain_loop_matches_graph_path(gpu):
    shots = [materialize(s) for s in mixed_specs(3, seed=5)]
    out_b, _ = serve(shots, max_active=3, steps_per_round=20)
    out_p, stats = serve(shots, plain_loop=True)
    assert stats["done"] == 3 and stats["failed"] == 0
    for (sb, eb, _), (sp, ep, _) in zip(out_b, out_p):
        assert np.array_equal(sb, sp) and np.array_equal(eb, ep)


def test_unstable_shot_fails_alone(gpu):
    specs = mixed_specs(4, seed=21)
    shots = [materialize(s) for s in specs]
    cfg, model = shots[1]
    shots[1] = (cfg.replace(dt=4.0 * cfg.dt, nstep=400), model)  # Courant 2: grows without bound
    with Server(max_active=4, steps_per_round=10) as srv:
        tickets = [srv.submit(c, *m) for c, m in shots]
        res = [srv.result(t) for t in tickets]
        stats = srv.stats()
    assert res[1][2]["status"] == -3  # ELA3D_E_UNSTABLE
    assert stats["failed"] == 1
    for q in (0, 2, 3):
        seis, en, info = res[q]
        assert info["status"] == 0
        rs, re = alone(*shots[q])
        assert np.array_equal(seis, rs)


def test_budget_rejection_and_tickets(gpu):
    cfg, model = materialize(mixed_specs(1)[0])
    need = cfg.plan()["device_bytes"]
    with Server(hbm_budget_bytes=need // 2) as srv:
        with pytest.raises(Ela3dError) as e:
            srv.submit(cfg, *model)
        assert e.value.code == -8  # ELA3D_E_BUDGET
        with pytest.raises(Ela3dError) as e:
            srv.poll(12345)
        assert e.value.code == -9  # ELA3D_E_NOTFOUND
    with Server() as srv:
        t = srv.submit(cfg, *model)
        srv.result(t)  # releases
        with pytest.raises(Ela3dError):
            srv.poll(t)
        cp, cs, rho = (np.array(a, copy=True) for a in model)
        cs.reshape(-1)[7] = cp.reshape(-1)[7]  # cs == cp: no valid Lame parameters
        with pytest.raises(Ela3dError) as e:
            srv.submit(cfg, cp, cs, rho)
        assert e.value.code == -1 and "model cell 7" in str(e.value)
        assert srv.stats()["submitted"] == 1


def test_served_shot_vs_fp64_oracle(gpu):
    spec = [s for s in mixed_specs(12, seed=8) if s.layered][0]
    cfg, model = materialize(spec)
    out, _ = serve([(cfg, model)])
    seis, en, info = out[0]
    assert info["status"] == 0
    s64, e64, _ = oracle.run(cfg, model, dtype=np.float64, threads=16)
    err = rel_l2(seis, s64)
    print(f"served {cfg.nx}^3 layered shot: seismogram rel L2 vs FP64 {err:.3e}")
    assert err <= 1e-4
    assert np.allclose(en, e64, rtol=1e-4, atol=1e-12 * np.abs(e64).max())


def test_mixed_storage_shots_match_alone(gpu):
    """FP16- and FP64-storage shots served among FP32 ones: each keeps its own scale
    table and field sets, so every shot is bit-identical to its run alone."""
    from ela3d_amd import STORAGE_FP16, STORAGE_FP64
    kinds = [STORAGE_FP16, STORAGE_FP64, None]
    shots = []
    for q, spec in enumerate(mixed_specs(6, seed=21)):
        cfg, model = materialize(spec)
        if kinds[q % 3] is not None:
            cfg = cfg.replace(storage=kinds[q % 3])
        shots.append((cfg, model))
    ref = [alone(c, m) for c, m in shots]
    out, stats = serve(shots, max_active=3, steps_per_round=20)
    assert stats["done"] == len(shots) and stats["failed"] == 0
    for (seis, en, info), (rs, re), (cfg, _) in zip(out, ref, shots):
        assert info["status"] == 0
        assert np.array_equal(seis, rs), f"storage {cfg.storage} shot {cfg.nx}^3 differs in the mix"
        assert np.array_equal(en, re)
