"""Serving loop on the CPU oracle (small windows): admission against the HBM budget,
round schedule, the one-period CW criterion, latency bookkeeping, and isolation (every
result identical alone and in the mix)."""
import numpy as np

from lens_sensor_amd import serving
from oracle.engine import OracleEngine


def _factory(kind, dims, mode):
    return OracleEngine(kind, dims, mode)


def _small_stream():
    return [dict(id=0, arrival=0.0, kind='2d', dims=(130, 460), pol='s', bead_x_frac=0.5,
                 bead_r_um=0.3, kappa_aq=0.001),
            dict(id=1, arrival=0.01, kind='2d', dims=(150, 464), pol='p', bead_x_frac=0.4,
                 bead_r_um=0.25, kappa_aq=0.0),
            dict(id=2, arrival=0.02, kind='3d', dims=(40, 110, 22), bead_x_frac=0.5, bead_r_um=0.2)]


def test_request_stream_is_seeded_poisson_with_mixed_kinds():
    r = serving.request_stream(400, rate_hz=20.0, seed=1)
    t = np.array([q['arrival'] for q in r])
    assert np.all(np.diff(t) > 0)
    assert abs(np.mean(np.diff(t)) - 0.05) < 0.01                  # mean 20 req/s
    frac3 = np.mean([q['kind'] == '3d' for q in r])
    assert 0.12 < frac3 < 0.28
    assert r == serving.request_stream(400, rate_hz=20.0, seed=1)


def test_budget_admission_and_isolation_on_oracle():
    reqs = _small_stream()
    srv = serving.Server(engine_factory=_factory, batched=False, round_steps=96)
    recs, summ = srv.run(reqs, realtime=False)
    assert summ['requests'] == len(reqs)
    by_id = {r['id']: r for r in recs}
    # alone: each request in a server of its own
    for q in reqs:
        one, _ = serving.Server(engine_factory=_factory, batched=False, round_steps=96).run(
            [dict(q, arrival=0.0)], realtime=False)
        a, b = one[0], by_id[q['id']]
        assert a['steps'] == b['steps']
        assert np.array_equal([a['transmission'], a['change'], a['phasor_norm']],
                              [b['transmission'], b['change'], b['phasor_norm']], equal_nan=True)
    # a budget that holds only one request at a time serialises admission
    tiny = serving.Server(engine_factory=_factory, batched=False, round_steps=96, hbm_budget=1.0)
    recs1, _ = tiny.run(reqs, realtime=False)
    adm = sorted((r['admitted'], r['finished']) for r in recs1)
    for (a0, f0), (a1, f1) in zip(adm, adm[1:]):
        assert a1 >= f0 - 1e-9                                       # one at a time
