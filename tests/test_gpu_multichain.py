"""Several independent chains stepped together (nrgk_chains_step): each chain must end exactly where
it ends when it is stepped alone -- the chains share nothing but the device -- and so matches the
oracle where the single-chain path does.

Reference: the per-site loop of examples/fdmSpectrumSiam/main.cpp:86-95 (add_bath_site,
update_system_operator, saveCurrentData when higher energies were discarded, update_internal_state).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L, EPS, U, G = 4.0, -0.25, 0.5, 0.015


def _z_hoppings(nk, zs, ns):
    A = 0.5 * np.log(L) * (1 + 1 / L) / (1 - 1 / L)
    return [nk.z_chain(z, L, ns)[0] * A for z in zs]


def _thermo_alone(nk, hop, K, T):
    g = nk.NrgCore(nk.Site.spinhalf(EPS, U), nk.Site.spinhalf(0, 0), K)
    rec = nk.FdmRecorder(g, None)
    V = np.sqrt(2 * G / np.pi)
    g.add_bath_site([V, V], 1.0)
    g.update_internal_state()
    min_it = 0
    for n in range(len(hop)):
        g.add_bath_site([hop[n], hop[n]], 1.0 if n == 0 else np.sqrt(L))
        if g.checkHigherEnergyDiscarded():
            rec.record()
        else:
            min_it += 1
        g.update_internal_state()
    S, Cv = rec.thermodynamics(L, T, scale_offset=-1, first_cnt=min_it + 1)
    return S, Cv, g.all_eigenvalue.copy(), g.kept_counts()


@pytest.mark.parametrize("n_threads", [0, 3])
def test_z_chains_stepped_together_equal_chains_stepped_alone(nk, oracle_mod, n_threads):
    """Eight z-shifted chains (z = k/8) with thermodynamics recorders, stepped together with the
    default worker count and with 3 workers: entropy, specific heat, eigenvalues and kept counts of
    every chain bit-identical to the chain stepped alone; the z = 1 chain also against the oracle."""
    K, ns = 256, 16
    zs = [(k if k else 8) / 8 for k in range(8)]
    hops = _z_hoppings(nk, zs, ns)
    T = np.logspace(-6, 0, 25)
    gs = [nk.NrgCore(nk.Site.spinhalf(EPS, U), nk.Site.spinhalf(0, 0), K) for _ in zs]
    recs = [nk.FdmRecorder(g, None) for g in gs]
    V = np.sqrt(2 * G / np.pi)
    nk.chains_step(gs, [[V, V]] * len(gs), 1.0, n_threads=n_threads)
    min_it = np.zeros(len(gs), int)
    for n in range(ns):
        t = [[h[n], h[n]] for h in hops]
        done = nk.chains_step(gs, t, 1.0 if n == 0 else np.sqrt(L), recorders=recs, n_threads=n_threads)
        min_it += ~done
    for k, z in enumerate(zs):
        S, Cv = recs[k].thermodynamics(L, T, scale_offset=-1, first_cnt=int(min_it[k]) + 1)
        S1, Cv1, e1, kc1 = _thermo_alone(nk, hops[k], K, T)
        assert np.array_equal(S, S1) and np.array_equal(Cv, Cv1), z
        assert np.array_equal(gs[k].all_eigenvalue, e1) and np.array_equal(gs[k].kept_counts(), kc1), z
    k1 = zs.index(1.0)
    drv = oracle_mod.driver(imp_kind=0, eps=EPS, U=U, bath_kind=0, Lambda=L, first_V=V, nsites=ns, K=K)
    oS, _ = oracle_mod.run_thermo_hoppings(drv, hops[k1], T, 0)
    S, _ = recs[k1].thermodynamics(L, T, scale_offset=-1, first_cnt=int(min_it[k1]) + 1)
    m = T >= 1e-5
    assert np.abs(S - oS)[m].max() < 1e-8


def test_fdm_chains_with_observables_equal_chains_stepped_alone(nk):
    """Three SIAM chains with different U, each with its f-dagger chi observable and an FDM recorder,
    stepped together: the forward pass (eigenvalues, kept counts, recorded shells) bit-identical to
    each chain stepped alone; the spectral weights to 1e-12 of their largest bin -- the binning kernel
    adds into the bins with atomic adds, whose order (and so the last bit) changes from run to run
    for one chain alone as well."""
    Lw, K, ns, nb = 2.0, 128, 16, 300
    params = [(-0.15, 0.3), (-0.25, 0.5), (-0.1, 0.2)]
    gs = [nk.NrgCore(nk.Site.spinhalf(e, u), nk.Site.spinhalf(0, 0), K) for e, u in params]
    ops = [nk.SystemOperator(g, nk.Site.fdag_chi_spinhalf()) for g in gs]
    recs = [nk.FdmRecorder(g, op, nb, 1e-10, 10.0) for g, op in zip(gs, ops)]
    V = np.sqrt(2 * 0.01 / np.pi)
    nk.chains_step(gs, [[V, V]] * 3, 1.0, observables=[[o] for o in ops])
    min_it = np.zeros(3, int)
    for n in range(ns):
        t = nk.hopping(n, Lw)
        min_it += ~nk.chains_step(gs, [[t, t]] * 3, 1.0 if n == 0 else np.sqrt(Lw),
                                  observables=[[o] for o in ops], recorders=recs)
    for k, (e, u) in enumerate(params):
        pos, neg, tr, sm = recs[k].backward(Lw, scale_offset=-1, first_cnt=int(min_it[k]) + 1)
        g = nk.NrgCore(nk.Site.spinhalf(e, u), nk.Site.spinhalf(0, 0), K)
        op = nk.SystemOperator(g, nk.Site.fdag_chi_spinhalf())
        rec = nk.FdmRecorder(g, op, nb, 1e-10, 10.0)
        g.add_bath_site([V, V], 1.0)
        op.update()
        g.update_internal_state()
        m1 = 0
        for n in range(ns):
            t = nk.hopping(n, Lw)
            g.add_bath_site([t, t], 1.0 if n == 0 else np.sqrt(Lw))
            op.update()
            if g.checkHigherEnergyDiscarded():
                rec.record()
            else:
                m1 += 1
            g.update_internal_state()
        assert m1 == min_it[k]
        assert np.array_equal(gs[k].all_eigenvalue, g.all_eigenvalue), k
        assert np.array_equal(gs[k].kept_counts(), g.kept_counts()), k
        p1, n1, tr1, sm1 = rec.backward(Lw, scale_offset=-1, first_cnt=m1 + 1)
        top = max(np.abs(p1).max(), np.abs(n1).max())
        assert np.abs(pos - p1).max() <= 1e-12 * top and np.abs(neg - n1).max() <= 1e-12 * top, k
        np.testing.assert_allclose(tr, tr1, rtol=1e-12)
        np.testing.assert_allclose(sm, sm1, rtol=1e-12)


@pytest.mark.parametrize("max_workers", ["1", "2"])
def test_chains_step_when_worker_threads_cannot_start(nk, monkeypatch, max_workers):
    """Thread creation failing after k - 1 workers (NRGK_DEBUG_MAX_WORKERS=k; 1: none start) leaves
    the remaining chains to the workers that did start and the calling thread: five chains asked for
    five workers still all step, bit-identical to the chains stepped alone."""
    monkeypatch.setenv("NRGK_DEBUG_MAX_WORKERS", max_workers)
    K, ns = 128, 8
    zs = [0.2, 0.4, 0.6, 0.8, 1.0]
    hops = _z_hoppings(nk, zs, ns)
    gs = [nk.NrgCore(nk.Site.spinhalf(EPS, U), nk.Site.spinhalf(0, 0), K) for _ in zs]
    V = np.sqrt(2 * G / np.pi)
    nk.chains_step(gs, [[V, V]] * len(gs), 1.0, n_threads=5)
    for n in range(ns):
        nk.chains_step(gs, [[h[n], h[n]] for h in hops], 1.0 if n == 0 else np.sqrt(L), n_threads=5)
    monkeypatch.delenv("NRGK_DEBUG_MAX_WORKERS")
    for k, g in enumerate(gs):
        a = nk.NrgCore(nk.Site.spinhalf(EPS, U), nk.Site.spinhalf(0, 0), K)
        a.add_bath_site([V, V], 1.0)
        a.update_internal_state()
        for n in range(ns):
            a.add_bath_site([hops[k][n]] * 2, 1.0 if n == 0 else np.sqrt(L))
            a.update_internal_state()
        assert g.nrg_iterations_cnt == a.nrg_iterations_cnt
        assert np.array_equal(g.all_eigenvalue, a.all_eigenvalue) and np.array_equal(g.kept_counts(), a.kept_counts())


def test_chains_step_errors(nk):
    """A chain listed twice, an observable or recorder of another chain: rejected before any chain
    moves.  A chain whose own step fails (wrong hopping count) is named; the others complete."""
    K = 64
    a = nk.NrgCore(nk.Site.spinhalf(EPS, U), nk.Site.spinhalf(0, 0), K)
    b = nk.NrgCore(nk.Site.spinhalf(EPS, U), nk.Site.spinhalf(0, 0), K)
    oa = nk.SystemOperator(a, nk.Site.fdag_chi_spinhalf())
    ob = nk.SystemOperator(b, nk.Site.fdag_chi_spinhalf())
    with pytest.raises(nk.NrgkError, match="twice"):
        nk.chains_step([a, a], [[0.1, 0.1]] * 2, 1.0)
    with pytest.raises(nk.NrgkError, match="observable"):
        nk.chains_step([a, b], [[0.1, 0.1]] * 2, 1.0, observables=[[ob], [oa]])
    with pytest.raises(nk.NrgkError, match="recorder"):
        nk.chains_step([a, b], [[0.1, 0.1]] * 2, 1.0, recorders=[nk.FdmRecorder(b, None), None])
    a0, b0 = a.nrg_iterations_cnt, b.nrg_iterations_cnt
    c = nk.NrgCore(nk.Site.spinhalf(EPS, U).duplicate_ops(), nk.Site.two_channel(), K)  # 4 operators
    c0 = c.nrg_iterations_cnt
    with pytest.raises(nk.NrgkError, match="chain 1"):  # 2 hoppings for 4 operators (nrgcore.hpp:148-152)
        nk.chains_step([a, c], [[0.1, 0.1]] * 2, 1.0)
    assert a.nrg_iterations_cnt == a0 + 1 and c.nrg_iterations_cnt == c0 and b.nrg_iterations_cnt == b0
    assert len(nk.chains_step([], np.zeros((0, 2)), [])) == 0


def test_chains_with_wide_blocks_stepped_together(nk, monkeypatch):
    """Three two-channel chains whose wide blocks (limit lowered to 40 rows) go to their own
    rocSOLVER side streams, stepped together: each bit-identical to the chain stepped alone."""
    monkeypatch.setenv("NRGK_MIXED_MAX", "40")
    K, Lw, ns = 128, 3.0, 4
    params = [(-0.2, 0.4), (-0.25, 0.5), (-0.15, 0.3)]

    def make(e, u):
        return nk.NrgCore(nk.Site.spinhalf(e, u).duplicate_ops(), nk.Site.two_channel(), K)

    V = np.sqrt(2 * 0.01 / np.pi)
    ts = [V] + [nk.hopping(n, Lw) for n in range(ns)]
    rs = [1.0] + [1.0 if n == 0 else np.sqrt(Lw) for n in range(ns)]
    gs = [make(e, u) for e, u in params]
    for t, r in zip(ts, rs):
        nk.chains_step(gs, [[t] * 4] * 3, r)
    assert max(int(g.sectors()[1].max()) for g in gs) > 40
    for k, (e, u) in enumerate(params):
        g = make(e, u)
        for t, r in zip(ts, rs):
            g.add_bath_site([t] * 4, r)
            g.update_internal_state()
        assert np.array_equal(gs[k].all_eigenvalue, g.all_eigenvalue), k
        assert np.array_equal(gs[k].kept_counts(), g.kept_counts()), k
