"""Serving loop (continuous batching) on the GPU: a world's result is the same alone and in the
mix, bit for bit, while other worlds arrive, run and retire around it, and equal to the oracle
stepping that world alone; admission waits for room under the HBM budget; a request larger than
the batch is rejected without blocking the queue; a world that fails on the device is evicted
while the others go on unchanged."""
import math

import numpy as np
import pytest

from helpers import oracle, pkg, scenes, state_mismatch

pytestmark = pytest.mark.gpu


def _serve():
    from importlib import import_module
    return import_module("godot-avbd_amd.serve")


@pytest.mark.parametrize("overlap", [True, False])
def test_same_alone_and_in_the_mix(overlap):
    serve = _serve()
    params = pkg.default_params(iterations=10)
    srv = serve.Server(params, 1.5e9, max_worlds=64, joints_per_body=0.5, overlap=overlap)
    srv.keep_final = True
    rng = np.random.default_rng(11)
    background = []
    for k in range(8):
        if k % 3 == 2:
            b, j = scenes.chains(n_chains=4 + k, links=16, seed=100 + k)
        else:
            b, j = scenes.pile(int(rng.integers(60, 700)), seed=200 + k, ground=80.0)
        background.append(serve.Request(k, b, j, int(rng.integers(20, 90)), 0.0))
    targets = [serve.Request(100, *scenes.pile(300, seed=7, ground=80.0), 60, 0.0),
               serve.Request(101, *scenes.chains(n_chains=6, links=16, seed=9), 50, 0.0)]
    for r in background[:5]:
        srv.submit(r)
    for tick in range(140):
        if tick == 5:
            srv.submit(targets[0])
        if tick == 12:
            for r in background[5:]:
                srv.submit(r)
            srv.submit(targets[1])
        srv.tick()
    assert not srv.resident and not srv.waiting
    finals = {req.rid: st for req, st in srv.finished}
    assert len(finals) == len(background) + len(targets)
    for t in targets:
        alone = serve.run_alone(params, t.bodies, t.joints, t.horizon)
        bad = state_mismatch(finals[t.rid], alone)
        assert not bad, (t.rid, bad)
        bad = state_mismatch(finals[t.rid], _oracle_alone(params, t))
        assert not bad, (t.rid, "oracle", bad)
    srv.close()


def _oracle_alone(params, req):
    o = oracle.OracleWorld(params)
    o.upload(req.bodies, req.joints)
    o.step(req.horizon)
    return o.download()


def _request_stream(serve, rng, n, first_id=0):
    """C5-shaped requests: bodies log-uniform in [16, 16384], 70 % random piles and 30 % jointed
    chains, short horizons."""
    out = []
    for k in range(n):
        nb = int(round(math.exp(rng.uniform(math.log(16), math.log(16384)))))
        horizon = int(rng.integers(3, 16))
        if rng.uniform() < 0.7:
            b, j = scenes.pile(nb, seed=300 + first_id + k, pitch=4.0, ground=max(40.0, 8.0 * math.sqrt(nb)))
        else:
            b, j = scenes.chains(n_chains=max(1, nb // 33), links=32, seed=400 + first_id + k)
        out.append(serve.Request(first_id + k, b, j, horizon, 0.0))
    return out


def test_poisson_stream_matches_the_oracle():
    """Poisson arrivals of C5-shaped requests through the serving loop (seeded), with one 9000-body
    pile and one jointed-chain world among them: each of the checked worlds ends bit-identical to
    the oracle stepping it alone."""
    serve = _serve()
    params = pkg.default_params(iterations=10)
    rng = np.random.default_rng(55)
    srv = serve.Server(params, 6e9, max_worlds=256, joints_per_body=0.5)
    srv.keep_final = True
    stream = _request_stream(serve, rng, 30)
    big = serve.Request(1000, *scenes.pile(9000, seed=77, pitch=4.0, ground=800.0), 12, 0.0)
    chain = serve.Request(1001, *scenes.chains(n_chains=20, links=32, seed=78), 14, 0.0)
    stream[4:4] = [big]
    stream[11:11] = [chain]
    pending = list(stream)
    tick = 0
    while pending or srv.waiting or srv.resident:
        for _ in range(rng.poisson(1.5) if pending else 0):
            if pending:
                assert srv.submit(pending.pop(0))
        srv.tick()
        tick += 1
        assert tick < 2000
    assert not srv.rejected and not srv.evicted
    finals = {req.rid: st for req, st in srv.finished}
    assert len(finals) == len(stream)
    checked = [big, chain] + [r for r in stream if len(r.bodies["mass"]) < 3000][:4]
    for t in checked:
        bad = state_mismatch(finals[t.rid], _oracle_alone(params, t))
        assert not bad, (t.rid, len(t.bodies["mass"]), bad)
    srv.close()


def test_oversized_request_is_rejected_and_does_not_block():
    serve = _serve()
    params = pkg.default_params(iterations=10)
    srv = serve.Server(params, serve.footprint(params, 4, 600, 0) * 1.05, max_worlds=4, joints_per_body=0.0)
    huge = serve.Request(0, *scenes.pile(srv.max_bodies + 50, seed=1, ground=300.0), 5, 0.0)
    small = serve.Request(1, *scenes.pile(100, seed=2, ground=60.0), 4, 0.0)
    assert srv.submit(huge) is False and srv.rejected == [huge]
    assert srv.submit(small)
    steps = sum(1 for _ in range(6) if srv.tick() > 0)
    assert steps == 4 and not srv.waiting and not srv.resident
    srv.close()


def test_failed_worlds_are_evicted_and_the_rest_equal_the_oracle():
    """One world's state is not finite (a NaN velocity), another's manifolds overflow its slots
    (40 boxes piled into one place); both are evicted after their first step while a pile and a
    chain world beside them step on and end equal to the oracle."""
    serve = _serve()
    params = pkg.default_params(iterations=10)
    srv = serve.Server(params, 1.0e9, max_worlds=16, joints_per_body=0.5)
    srv.keep_final = True
    nan_b, nan_j = scenes.pile(30, seed=5, ground=40.0)
    nan_b["velocity"] = nan_b["velocity"].copy()
    nan_b["velocity"][3, 1] = np.nan
    s = scenes.SceneBuilder()
    s.box((40, 40, 1), 0.0, 0.5, (0, 0, 0))
    for k in range(40):
        s.box((1, 1, 1), 1.0, 0.5, (0.001 * k, 0.0, 1.0 + 0.001 * k))
    over_b, over_j = s.build()
    keep = [serve.Request(0, *scenes.pile(250, seed=6, ground=60.0), 30, 0.0),
            serve.Request(1, *scenes.chains(n_chains=5, links=16, seed=7), 30, 0.0)]
    bad = [serve.Request(2, nan_b, nan_j, 30, 0.0), serve.Request(3, over_b, over_j, 30, 0.0)]
    for r in (keep[0], bad[0], keep[1], bad[1]):
        assert srv.submit(r)
    for _ in range(40):
        srv.tick()
    assert not srv.resident
    why = {req.rid: (reason, done) for req, reason, done in srv.evicted}
    assert why[2][0] & pkg.abi.AVBD_WORLD_NONFINITE and why[2][1] == 0
    assert why[3][0] & pkg.abi.AVBD_WORLD_SLOTS and why[3][1] == 0
    finals = {req.rid: st for req, st in srv.finished}
    for t in keep:
        bad_fields = state_mismatch(finals[t.rid], _oracle_alone(params, t))
        assert not bad_fields, (t.rid, bad_fields)
    srv.close()


def test_rejected_world_gives_its_ranges_back():
    """A world the batch rejects (an unknown body shape) takes no capacity: a world of the full
    body capacity still fits afterwards."""
    params = pkg.default_params(iterations=10)
    b, j = scenes.pile(200, seed=4, ground=60.0)
    batch = pkg.Batch(params, 2, len(b["mass"]), 0)
    bad = dict(b)
    bad["shape"] = np.full(len(b["mass"]), 7, np.uint8)
    with pytest.raises(pkg.abi.AvbdError):
        batch.add(bad, j)
    assert batch.fits(len(b["mass"]))[0]
    batch.add(b, j)
    batch.step()
    batch.close()


@pytest.mark.parametrize("overlap", [True, False])
def test_admission_waits_for_budget(overlap):
    """A budget that holds about one large world: the second waits until the first retires,
    then is admitted at the next tick and steps its horizon."""
    serve = _serve()
    params = pkg.default_params(iterations=10)
    b, j = scenes.pile(2000, seed=3, ground=200.0)
    need = serve.footprint(params, 4, 2000, 0)
    srv = serve.Server(params, need * 1.3, max_worlds=4, joints_per_body=0.0, overlap=overlap)
    first, second = serve.Request(0, b, j, 10, 0.0), serve.Request(1, b, j, 5, 0.0)
    srv.submit(first)
    srv.submit(second)
    srv.tick()
    assert len(srv.resident) == 1 and len(srv.waiting) == 1
    while first.done < first.horizon:
        assert second.wid < 0 or second.done == 0
        srv.tick()
    assert first.wid not in srv.resident and len(srv.waiting) == 1
    srv.tick()                                               # room again: admitted
    assert len(srv.waiting) == 0 and len(srv.resident) == 1
    while srv.resident:
        srv.tick()
    assert second.done == second.horizon
    srv.close()


def test_async_admission_during_a_step():
    """Worlds added while a step is in flight (second stream, pinned copies) are stepped from the
    next launch and end equal to the oracle; the caller's arrays can change as soon as the add
    returns; other calls are refused until the step is waited for."""
    params = pkg.default_params(iterations=10)
    resident = [scenes.pile(600, seed=41, ground=80.0), scenes.chains(n_chains=6, links=16, seed=42)]
    batch = pkg.Batch(params, 8, 4000, 400)
    for bb, jj in resident:
        batch.add(bb, jj)
    for _ in range(3):
        batch.step()
    late = [scenes.pile(400, seed=43, ground=60.0), scenes.chains(n_chains=4, links=16, seed=44)]
    kept = [({k: v.copy() for k, v in bb.items()}, None if jj is None else {k: v.copy() for k, v in jj.items()})
            for bb, jj in late]
    batch.step_launch()
    ids = [batch.add(bb, jj, overlap=True) for bb, jj in late]
    for bb, _ in late:                      # the library copied them during the call
        bb["position"][:] = 1e6
        bb["velocity"][:] = np.nan
    with pytest.raises(pkg.abi.AvbdError):
        batch.download(ids[0])
    with pytest.raises(pkg.abi.AvbdError):
        batch.add(*kept[0])                 # only asynchronous additions while in flight
    batch.step_wait()
    for _ in range(12):
        batch.step()
    for wid, (bb, jj) in zip(ids, kept):
        o = oracle.OracleWorld(params)
        o.upload(bb, jj)
        o.step(12)
        assert not state_mismatch(batch.download(wid), o.download()), wid
    batch.close()
