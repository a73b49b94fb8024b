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


def test_diverged_and_nan_cases_are_evicted_without_perturbing_the_batch():
    """A case seeded with a NaN (its pressure solve reports NaN, reason -9)
    and a case whose CG may take one iteration (reason -3) are evicted as
    failed after their first step; the healthy request beside them finishes
    with the result it has alone."""
    import dataclasses
    from tpls_amd.case import Case
    from tpls_amd.config import CaseParams
    from tpls_amd.initial import make_initial
    import tpls_amd.serve as S
    good = CaseParams(nx=64, ny=16, nz=32, Re=200.0, mu_plus=3.0, seed=31, ksp_pc="mg")
    nan = CaseParams(nx=32, ny=16, nz=24, seed=32)
    capped = dataclasses.replace(CaseParams(nx=48, ny=8, nz=16, seed=33), ksp_maxits=1)
    reqs = [Request(rid=0, params=good, steps=4, arrival=0.0, keep=("u2", "phi2", "pres")),
            Request(rid=1, params=nan, steps=4, arrival=0.0),
            Request(rid=2, params=capped, steps=4, arrival=0.0)]
    real = S.Server._admit

    def admit(self, r, now):                 # seed the NaN into the second request's u
        real(self, r, now)
        if r.rid == 1:
            c = self.active[-1][1]
            init = list(make_initial(r.params))
            init[0] = init[0].copy(order="F")
            init[0][5, 3, 4] = np.nan
            c.load(*init)
    srv = Server(1e12, graphs=True)
    srv._admit = admit.__get__(srv)
    rep = srv.run(reqs)
    print("[serve failures]", rep, [(r.rid, r.status, r.error) for r in reqs], flush=True)
    assert rep["requests_finished"] == 1 and rep["requests_failed"] == 2
    assert reqs[0].status == "done"
    assert reqs[1].status == "failed" and "-9" in reqs[1].error
    assert reqs[2].status == "failed" and "-3" in reqs[2].error
    assert srv.used == 0
    with Case(good, precision=reqs[0].precision, diagnostics=False, graphs=True, tiles="auto") as c:
        c.load(*make_initial(good))
        c.step(4)
        for f in reqs[0].keep:
            assert np.array_equal(c.field(f), reqs[0].result[f]), f
