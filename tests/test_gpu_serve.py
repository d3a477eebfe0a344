"""Serving loop (continuous batching) on the GPU: a world's result is the same alone and in the
mix, bit for bit, while other worlds arrive, run and retire around it; admission waits for
room under the HBM budget."""
import numpy as np
import pytest

from helpers import pkg, scenes, state_mismatch

pytestmark = pytest.mark.gpu


def _serve():
    from importlib import import_module
    return import_module("godot-avbd_amd.serve")


def test_same_alone_and_in_the_mix():
    serve = _serve()
    params = pkg.default_params(iterations=10)
    srv = serve.Server(params, 1.5e9, max_worlds=64, joints_per_body=0.5)
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
    srv.close()


def test_admission_waits_for_budget():
    """A budget that holds about one large world: the second waits until the first retires."""
    serve = _serve()
    params = pkg.default_params(iterations=10)
    b, j = scenes.pile(2000, seed=3, ground=200.0)
    need = serve.footprint(params, 4, 2000, 0)
    srv = serve.Server(params, need * 1.3, max_worlds=4, joints_per_body=0.0)
    first, second = serve.Request(0, b, j, 10, 0.0), serve.Request(1, b, j, 5, 0.0)
    srv.submit(first)
    srv.submit(second)
    srv.tick()
    assert len(srv.resident) == 1 and len(srv.waiting) == 1
    for _ in range(10):
        srv.tick()
    assert len(srv.waiting) == 0 and len(srv.resident) == 1   # admitted once the first retired
    for _ in range(5):
        srv.tick()
    assert not srv.resident
    srv.close()
