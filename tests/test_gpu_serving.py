"""The request loop (nrgk_server_*) on a mix of the serving workload's requests: SIAM spectral
functions of different U, eps, Gamma, Lambda, K and chain lengths, some arriving while others run.

  * every request's A(omega) in the mix equals the same request run alone: the forward pass is
    bit-identical, and the FDM bins are summed with atomics whose order varies, so A(omega) and the
    total weight agree to 1e-12 relative;
  * one request against the oracle (binned weights broadened on the same grid, 1e-6 of the maximum);
  * a request whose chain fails (a NaN hopping injected at one site) fails alone;
  * with a device-memory budget, the running requests' footprints never sum past it, fewer requests
    run at once than without it, and the results are unchanged."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OMEGA = np.logspace(-8, 0.5, 240)


def _requests(seed=20261203, n=8):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        U = rng.uniform(0.1, 1.0)
        out.append(dict(U=U, eps=-U / 2 + rng.uniform(-0.1, 0.1),
                        G=float(np.exp(rng.uniform(np.log(0.005), np.log(0.05)))),
                        L=float(rng.choice([2.0, 2.5, 3.0])), K=int(rng.choice([512, 1024])),
                        n_sites=int(rng.integers(18, 29))))
    return out


def _submit(nk, srv, r, fail_at_site=-1):
    return srv.submit(nk.Site.spinhalf(r["eps"], r["U"]), nk.Site.spinhalf(0, 0), nk.Site.fdag_chi_spinhalf(),
                      r["K"], r["L"], float(np.sqrt(2 * r["G"] / np.pi)), r["n_sites"], 2000, OMEGA, 0.6,
                      fail_at_site)


@pytest.fixture(scope="module")
def alone(nk):
    """each request run by itself (one slot, submitted after the previous one finished)"""
    srv = nk.Server(n_slots=1)
    res = []
    for r in _requests():
        res.append(srv.wait(_submit(nk, srv, r)))
    srv.close()
    return res


def _same(a, b):
    (ap, an, ia), (bp, bn, ib) = a, b
    assert ia["iterations"] == ib["iterations"] and ia["shells"] == ib["shells"]
    assert abs(ia["total_weight"] - ib["total_weight"]) <= 1e-12 * ib["total_weight"]
    scale = max(np.abs(bp).max(), np.abs(bn).max())
    assert np.abs(ap - bp).max() <= 1e-12 * scale and np.abs(an - bn).max() <= 1e-12 * scale


def test_mix_equals_alone_and_failure_is_isolated(nk, alone):
    reqs = _requests()
    srv = nk.Server(n_slots=4)
    ids = [_submit(nk, srv, r) for r in reqs[:3]]
    bad = _submit(nk, srv, reqs[3], fail_at_site=4)
    while not any(srv.done(i) for i in ids):  # the rest arrive while the first ones run
        time.sleep(0.01)
    ids += [_submit(nk, srv, r) for r in reqs[3:]]
    out = [srv.wait(i) for i in ids]
    with pytest.raises(nk.NrgkError, match="non-finite hopping"):
        srv.wait(bad)
    st = srv.stats()
    srv.close()
    assert st["running"] == 0 and st["queued"] == 0 and 2 <= st["peak_running"] <= 4
    for o, a in zip(out, alone):
        _same(o, a)


def test_budget_limits_concurrency(nk, alone):
    """the admitted footprints never sum past the budget; with a budget of two and a half of the
    largest footprint, at most as many run at once as the smallest footprints fit in it, and fewer
    than all of them"""
    reqs = _requests()
    fps = [a[2]["bytes"] for a in alone]
    budget = 2 * max(fps) + max(fps) // 2
    fit = int(np.searchsorted(np.cumsum(sorted(fps)), budget, side="right"))
    assert 2 <= fit < len(reqs)
    srv = nk.Server(n_slots=8, hbm_budget=budget)
    ids = [_submit(nk, srv, r) for r in reqs]
    out = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["peak_reserved"] <= budget and 1 <= st["peak_running"] <= fit
    for o, a in zip(out, alone):
        _same(o, a)


def test_one_request_matches_oracle(nk, oracle_mod, alone):
    r = _requests()[0]
    ap, an, info = alone[0]
    drv = oracle_mod.driver(imp_kind=0, eps=r["eps"], U=r["U"], bath_kind=0, Lambda=r["L"],
                            first_V=float(np.sqrt(2 * r["G"] / np.pi)), nsites=r["n_sites"], K=r["K"])
    pos, neg, _, _ = oracle_mod.run_fdm(drv, 2000)
    dele = (2000 - 1.0) / np.log(10 / 1e-10)
    x = 1e-10 * np.exp(np.arange(2000) / dele)
    op = oracle_mod.broaden(x, pos.sum(0), OMEGA, 0.6)
    on = oracle_mod.broaden(x, neg.sum(0), OMEGA, 0.6)
    scale = max(np.abs(op).max(), np.abs(on).max())
    assert np.abs(ap - op).max() <= 1e-6 * scale and np.abs(an - on).max() <= 1e-6 * scale
    assert abs(info["total_weight"] - (pos.sum() + neg.sum())) <= 1e-9 * info["total_weight"]
