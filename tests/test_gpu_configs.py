"""GPU parity at the benchmark workloads' own shapes, and on the checks that pin the oracle.

  * two-channel chain (16-state sites) against exact diagonalisation, iterations 0-1;
  * FDM complete-basis sum rule on the device (signed weights of every shell add up to <{B, B+}>);
  * the SIAM FDM spectrum workload (U = 0.3, eps = -0.15, Gamma = 0.01, Lambda = 2, 2000 bins) at
    K = 1024 against the oracle;
  * the two-channel workload at its real shape (Lambda = 3, K = 4096: blocks up to ~4200 rows, the
    wide-block branch of the eigensolver), every block's eigenvalues against LAPACK on the same H, and
    over its whole 50-site flow.

Tolerances as stated in tests/test_gpu_chain.py; the sum rule to 1e-10."""
import os
import time

import numpy as np
import pytest

from test_gpu_chain import _gpu_fdm, _jw, _state_from_oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_two_channel_chain_equals_exact_diagonalisation(nk):
    """examples/twoChannelSiam/main.cpp form, untruncated (K above 1024): the 64 and 1024 states of
    iterations 0 and 1 against the exact spectrum of the Jordan-Wigner chain (orbitals d_up, d_dn, then
    four per site: channel 1 up/down, channel 2 up/down), with unequal channel couplings
    {V_L, V_L, V_R, V_R} at the impurity."""
    eps, U, L = -0.2, 0.4, 3.0
    t0 = [0.11, 0.11, 0.05, 0.05]
    t1 = nk.hopping(0, L)
    g = nk.NrgCore(nk.Site.spinhalf(eps, U).duplicate_ops(), nk.Site.two_channel(), 100000)
    for n_sites, (t, r) in enumerate(((t0, 1.0), ([t1] * 4, 1.0)), start=1):
        g.add_bath_site(t, r)
        g.update_internal_state()
        n_orb = 2 + 4 * n_sites
        f = _jw(n_orb)
        nd = [f[s] @ f[s].T for s in (0, 1)]
        H = eps * (nd[0] + nd[1]) + U * nd[0] @ nd[1]
        for p in range(4):
            hop = f[p % 2] @ f[2 + p].T
            H = H + t0[p] * (hop + hop.T)
        if n_sites == 2:
            for p in range(4):
                hop = f[2 + p] @ f[6 + p].T
                H = H + t1 * (hop + hop.T)
        ed = np.linalg.eigvalsh(H)
        got = np.sort(g.all_eigenvalue)
        assert len(got) == len(ed)
        assert np.abs(got - ed).max() < 1e-10, (n_sites, np.abs(got - ed).max())


@pytest.mark.parametrize("K,nsites", [(256, 20), (1024, 24)])
def test_fdm_complete_basis_sum_rule_on_device(nk, K, nsites):
    """The signed weights of every recorded shell (before |.|, the 1e-20 cut and the bin range) add up
    to <{d, d+}> = 1 per spin.  K is a power of the bath dimension, so every truncating iteration is
    recorded (checkHigherEnergyDiscarded, nrgcore.hpp:384-386)."""
    _say("fdm sum rule K=%d nsites=%d" % (K, nsites))
    (_, _, tr, _), _, sg = _gpu_fdm(nk, -0.15, 0.3, 0.01, 2.0, K, nsites, 400, signed=True)
    np.testing.assert_allclose(tr, 1.0, atol=1e-10)
    np.testing.assert_allclose(sg, 1.0, atol=1e-10)


def test_fdm_spectrum_workload_at_K1024_matches_oracle(nk, oracle_mod):
    """The SIAM FDM workload's parameters (U = 0.3, eps = -0.15, Gamma = 0.01, Lambda = 2, 2000 bins
    over [1e-10, 10]) at K = 1024 and 40 sites: binned weights, per-shell sums and signed totals against
    the oracle; broadened spectra (b = 0.6) to 1e-6 of their maximum."""
    eps, U, G, L, K, ns, nb = -0.15, 0.3, 0.01, 2.0, 1024, 40, 2000
    _say("fdm workload K=%d nsites=%d bins=%d" % (K, ns, nb))
    t0 = time.time()
    (pos, neg, tr, sums), x, sg = _gpu_fdm(nk, eps, U, G, L, K, ns, nb, signed=True)
    _say("  gpu %.1f s" % (time.time() - t0))
    drv = oracle_mod.driver(imp_kind=0, eps=eps, U=U, bath_kind=0, Lambda=L, first_V=np.sqrt(2 * G / np.pi),
                            nsites=ns, K=K)
    opos, oneg, otr, osums, osg = oracle_mod.run_fdm(drv, nb, signed=True)
    _say("  oracle done %.1f s" % (time.time() - t0))
    assert len(tr) == len(otr)
    np.testing.assert_allclose(tr, 1.0, atol=1e-10)
    np.testing.assert_allclose(sums, osums, rtol=1e-8)
    np.testing.assert_allclose(sg, osg, atol=1e-10)
    np.testing.assert_allclose(sg, 1.0, atol=1e-10)
    for p in range(2):
        assert abs(pos[p].sum() + neg[p].sum() - (opos[p].sum() + oneg[p].sum())) <= 1e-9
    om = np.concatenate([-np.logspace(0, -9, 300), np.logspace(-9, 0, 300)])
    for p in range(2):
        for a, b in ((pos[p], opos[p]), (neg[p], oneg[p])):
            ag, ao = nk.broaden(x, a, np.abs(om), 0.6), oracle_mod.broaden(x, b, np.abs(om), 0.6)
            assert np.abs(ag - ao).max() <= 1e-6 * max(1e-12, np.abs(ao).max())


def _truncate(eigs, K, errbar=1e-8):
    """kept counts of the reference's rule (nrgcore.hpp:269-327) from per-sector eigenvalues"""
    allE = np.sort(np.concatenate([e for e in eigs if len(e)]))
    if len(allE) <= K:
        return np.array([len(e) for e in eigs])
    cut = allE[K] + errbar
    return np.array([int(np.sum(e < cut)) for e in eigs])


def test_two_channel_workload_at_full_shape(nk, oracle_mod):
    """The two-channel workload (U = 0.4, eps = -0.2, Gamma_1 = Gamma_2 = 0.01, Lambda = 3, K = 4096) at
    its real shape: iteration 2 (16 384 states, blocks up to 1 225) in lockstep with the oracle (sector
    labels and dims, assembled Hamiltonians and kept counts bit-exact), then iterations 3-4 (66 048
    states, 28 blocks wider than 768, up to ~4 200) free-running.  Every step: each block's eigenvalues
    within 1e-10 of LAPACK on the same H, the two widest blocks' eigenvectors orthonormal and their
    residual ||H U - U diag(E)|| within 1e-10, kept counts equal to the truncation rule applied to the
    LAPACK eigenvalues."""
    eps, U, L, K = -0.2, 0.4, 3.0, 4096
    V = np.sqrt(2 * 0.01 / np.pi)
    g = nk.NrgCore(nk.Site.spinhalf(eps, U).duplicate_ops(), nk.Site.two_channel(), K)
    g.keep_hamiltonians(True)
    o = oracle_mod.Chain(oracle_mod.Model("spinhalf", eps=eps, U=U, dup_ops=True), oracle_mod.Model("two_channel"), K,
                         keep_h=True)
    ts = [V] + [nk.hopping(n, L) for n in range(4)]
    rs = [1.0, 1.0, np.sqrt(L), np.sqrt(L), np.sqrt(L)]
    for it, (t, r) in enumerate(zip(ts, rs)):
        t0 = time.time()
        lock = it <= 2
        if lock and it > 0:
            g.set_state(_state_from_oracle(nk, o), o.cnt)
        g.add_bath_site([t] * 4, r)
        qn, dims = g.sectors()
        if lock:
            o.add_bath_site([t] * 4, r)
            assert np.array_equal(qn, o.cur_qn()) and np.array_equal(dims, o.sector_dims())
        eq = g.eigenvaluesQ
        lap = []
        wide = sorted(range(len(dims)), key=lambda s: -dims[s])[:2]
        for s, n in enumerate(dims):
            if n == 0:
                lap.append(np.zeros(0))
                continue
            H = g.block_h(s)
            if lock:
                assert np.array_equal(H, o.block_h(s, n)), (it, s)
            w = np.linalg.eigvalsh(H)
            lap.append(w)
            assert np.abs(w - eq[s]).max() <= 1e-10, (it, s, n, np.abs(w - eq[s]).max())
            if s in wide and n > 1:
                Us = g.current_hamiltonQ(s)
                orth = np.abs(Us.T @ Us - np.eye(n)).max()
                res = np.abs(H @ Us - Us * eq[s]).max()
                print("iteration %d sector %d n %d: orthogonality %.2e residual %.2e" % (it, s, n, orth, res),
                      flush=True)
                assert orth <= 1e-10 and res <= 1e-10, (it, s, n, orth, res)
        g.update_internal_state()
        kept = g.kept_counts()
        assert np.array_equal(kept, _truncate(lap, K)), it
        if lock:
            o.update_internal_state()
            assert np.array_equal(kept, o.kept_counts()), it
        _say("two-channel K=4096 iteration %d: %d states, widest block %d, %.1f s" %
             (it, int(np.sum(dims)), int(np.max(dims)), time.time() - t0))


def test_two_channel_workload_whole_flow(nk):
    """The two-channel workload over all 50 Wilson sites (the benchmark times iterations 3-10 of it):
    every iteration's eigenvalues finite and its kept counts equal to the truncation rule applied to
    its own eigenvalues; at iterations 15, 30 and 49 (near the fixed point, where the clusters of close
    eigenvalues are largest) each block's eigenvalues within 1e-10 of LAPACK on the same H and the
    widest block's eigenvectors orthonormal with residual within 1e-10."""
    eps, U, L, K, N = -0.2, 0.4, 3.0, 4096, 50
    V = np.sqrt(2 * 0.01 / np.pi)
    g = nk.NrgCore(nk.Site.spinhalf(eps, U).duplicate_ops(), nk.Site.two_channel(), K)
    t0 = time.time()
    for it in range(N):
        t = V if it == 0 else nk.hopping(it - 1, L)
        r = 1.0 if it <= 1 else np.sqrt(L)
        check = it in (15, 30, N - 1)
        g.keep_hamiltonians(check)
        g.add_bath_site([t] * 4, r)
        _, dims = g.sectors()
        eq = g.eigenvaluesQ
        assert all(np.all(np.isfinite(e)) for e in eq), it
        if check:
            widest = int(np.argmax(dims))
            worst = 0.0
            for s, n in enumerate(dims):
                if n == 0:
                    continue
                H = g.block_h(s)
                worst = max(worst, float(np.abs(np.linalg.eigvalsh(H) - eq[s]).max()))
                if s == widest:
                    Us = g.current_hamiltonQ(s)
                    orth = np.abs(Us.T @ Us - np.eye(n)).max()
                    res = np.abs(H @ Us - Us * eq[s]).max()
                    assert orth <= 1e-10 and res <= 1e-10, (it, s, n, orth, res)
            _say("two-channel K=4096 iteration %d: widest block %d, eigenvalue error %.2e, orthogonality %.2e, "
                 "residual %.2e" % (it, int(np.max(dims)), worst, orth, res))
            assert worst <= 1e-10, (it, worst)
        expect = _truncate([np.asarray(e, float) for e in eq], K)
        g.update_internal_state()
        assert np.array_equal(g.kept_counts(), expect), it
    _say("two-channel K=4096, %d iterations: %.1f s" % (N, time.time() - t0))
