"""Continuous-batching serving loop (tpls_amd/serve.py): requests that arrive
while others run are admitted as the HBM budget allows and join the batch;
each result equals the result of the same request run alone, bit for bit."""
import numpy as np
import pytest

from tpls_amd.serve import Request, Server, poisson_requests

pytestmark = pytest.mark.gpu

SMALL = ((64, 16, 32), (128, 32, 64))


def test_results_same_alone_and_in_the_mix():
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    reqs = poisson_requests(rate=40.0, duration=0.4, seed=3, big_fraction=0.0, steps=(3, 12), grids=SMALL)
    assert len(reqs) >= 6
    for r in reqs:
        r.keep = ("phi2", "u2", "w2")
    budget = 3.2 * max(Case.device_bytes(r.params, r.precision) for r in reqs)
    srv = Server(budget, graphs=True)
    rep = srv.run(reqs)
    print("[serve]", rep, flush=True)
    assert rep["requests_finished"] == len(reqs)
    assert rep["batch_size_max"] >= 2                      # requests overlapped
    assert rep["hbm_peak_gb"] * 1e9 <= budget
    for r in reqs:
        with Case(r.params, precision=r.precision, diagnostics=False) as c:
            c.load(*make_initial(r.params))
            c.step(r.steps)
            for f in r.keep:
                assert np.array_equal(c.field(f), r.result[f]), (r.rid, f)


def test_budget_is_respected_and_fifo():
    from tpls_amd.case import Case
    reqs = poisson_requests(rate=50.0, duration=0.3, seed=5, big_fraction=0.0, steps=(2, 4), grids=SMALL[:1])
    one = Case.device_bytes(reqs[0].params, reqs[0].precision)
    srv = Server(2.5 * one, graphs=False)                  # at most two cases at a time
    rep = srv.run(reqs)
    assert rep["batch_size_max"] <= 2
    admitted = [r.admitted for r in sorted(srv.done, key=lambda r: r.rid)]
    assert admitted == sorted(admitted)                    # admitted in arrival order


def test_request_larger_than_budget_is_refused():
    from tpls_amd.config import CaseParams
    r = Request(rid=0, params=CaseParams(nx=64, ny=16, nz=32), steps=1, arrival=0.0)
    with pytest.raises(ValueError):
        Server(1e6).run([r])
