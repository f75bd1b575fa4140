"""Randomized sweeps against the oracle and against constructed spectra (opt-in: NRGK_RUN_SWEEPS=1;
they take a few minutes, beyond what the default GPU suite should spend).

  * chains: SIAM with U in [0, 1.5], eps in [-1.2, 0.4] independent of U, Gamma in [1e-4, 0.1],
    Lambda in [1.8, 6], K in {64, 256, 512}, 17 free-running iterations; two-channel (K = 128) and
    Kondo (K = 256) chains: kept counts bit-exact, all_eigenvalue within 1e-8 of the oracle;
  * FDM: random SIAM FDM runs (K = 128, 16 sites, 400 bins) against the oracle's T = 0 FDM: shell
    count, total weight to 1e-9, per-shell sums to 1e-8, broadened spectra to 1e-6 of their maximum;
  * eigensolver: 40 seeds of the constructed-spectra batch and 20 random batches (sizes 1..899).
"""
import os

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("NRGK_RUN_SWEEPS") != "1", reason="opt-in sweep: NRGK_RUN_SWEEPS=1")]


def _free_run(g, o, ts, rs, nops):
    worst = 0.0
    for it, (t, r) in enumerate(zip(ts, rs)):
        g.add_bath_site([t] * nops, r)
        o.add_bath_site([t] * nops, r)
        g.update_internal_state()
        o.update_internal_state()
        assert np.array_equal(g.kept_counts(), o.kept_counts()), it
        a, b = g.all_eigenvalue, o.all_eigenvalue()
        assert len(a) == len(b), it
        worst = max(worst, float(np.abs(a - b).max()))
    return worst


def test_random_siam_chains(nk, oracle_mod):
    rng = np.random.default_rng(99)
    for c in range(14):
        U, eps = float(rng.uniform(0.0, 1.5)), float(rng.uniform(-1.2, 0.4))
        G, L, K = float(10 ** rng.uniform(-4, -1)), float(rng.uniform(1.8, 6.0)), int(rng.choice([64, 256, 512]))
        V = np.sqrt(2 * G / np.pi)
        g = nk.NrgCore(nk.Site.spinhalf(eps, U), nk.Site.spinhalf(0, 0), K)
        o = oracle_mod.Chain(oracle_mod.Model("spinhalf", eps=eps, U=U), oracle_mod.Model("spinhalf"), K)
        ts = [V] + [oracle_mod.hopping(n, L) for n in range(16)]
        rs = [1.0] + [1.0 if n == 0 else np.sqrt(L) for n in range(16)]
        assert _free_run(g, o, ts, rs, 2) < 1e-8, (c, U, eps, G, L, K)


def test_random_two_channel_and_kondo_chains(nk, oracle_mod):
    rng = np.random.default_rng(7)
    for _ in range(4):
        U, eps = float(rng.uniform(0.1, 1.0)), float(rng.uniform(-0.8, 0.0))
        G, L = float(10 ** rng.uniform(-3, -1.5)), float(rng.uniform(2.5, 4))
        V = np.sqrt(2 * G / np.pi)
        g = nk.NrgCore(nk.Site.spinhalf(eps, U).duplicate_ops(), nk.Site.two_channel(), 128)
        o = oracle_mod.Chain(oracle_mod.Model("spinhalf", eps=eps, U=U, dup_ops=True), oracle_mod.Model("two_channel"), 128)
        ts = [V] + [oracle_mod.hopping(n, L) for n in range(5)]
        rs = [1.0] + [1.0 if n == 0 else np.sqrt(L) for n in range(5)]
        assert _free_run(g, o, ts, rs, 4) < 1e-8, (U, eps, G, L)
    for _ in range(4):
        J, L = float(rng.uniform(0.05, 0.5)), float(rng.uniform(2.0, 4.0))
        g = nk.NrgCore(nk.Site.kondo(J, 0.5), nk.Site.spinhalf(0, 0), 256)
        o = oracle_mod.Chain(oracle_mod.Model("kondo", J=J, S=0.5), oracle_mod.Model("spinhalf"), 256)
        ts = [oracle_mod.hopping(n, L) for n in range(14)]
        assert _free_run(g, o, ts, [np.sqrt(L)] * 14, 2) < 1e-8, (J, L)


def test_random_fdm(nk, oracle_mod):
    from test_gpu_chain import _gpu_fdm
    rng = np.random.default_rng(31)
    omg = np.logspace(-8, 0, 200)
    for _ in range(6):
        U, eps = float(rng.uniform(0.1, 1.0)), float(rng.uniform(-0.8, 0.2))
        G, L = float(10 ** rng.uniform(-3, -1.3)), float(rng.uniform(2.0, 3.5))
        (pos, neg, tr, sums), x = _gpu_fdm(nk, eps, U, G, L, 128, 16, 400)
        drv = oracle_mod.driver(imp_kind=0, eps=eps, U=U, bath_kind=0, Lambda=L, first_V=np.sqrt(2 * G / np.pi),
                                nsites=16, K=128)
        opos, oneg, otr, osums = oracle_mod.run_fdm(drv, 400)
        assert len(tr) == len(otr)
        tot, otot = pos.sum() + neg.sum(), opos.sum() + oneg.sum()
        assert abs(tot - otot) <= 1e-9 * otot
        np.testing.assert_allclose(sums, osums, rtol=1e-8)
        for p in range(2):
            for a, b in ((pos[p], opos[p]), (neg[p], oneg[p])):
                ag, ao = nk.broaden(x, a, omg, 0.6), oracle_mod.broaden(x, b, omg, 0.6)
                assert np.abs(ag - ao).max() <= 1e-6 * max(1e-12, np.abs(ao).max())


def test_eigensolver_sweep(nk):
    import test_gpu_eigh as te
    for seed in range(10, 50):
        blocks, lams = te._case(seed)
        te._check(blocks, lams, nk.eigh_batched(blocks))
    rng = np.random.default_rng(123)
    kinds = ["degenerate", "clustered", "graded", "uniform", "threshold"]
    for _ in range(20):
        blocks, lams = [], []
        for n in rng.integers(1, 900, 10):
            lam = te._spectrum(rng, int(n), kinds[rng.integers(5)])
            Q = te._haar(rng, int(n))
            H = (Q * lam) @ Q.T
            blocks.append(0.5 * (H + H.T))
            lams.append(lam)
        te._check(blocks, lams, nk.eigh_batched(blocks))
