"""The FP64 oracle against what pins it: the reference's own HDF5 outputs (converted to
tests/golden/*.npz by tests/golden/make_fixtures.py), closed forms and brute-force exact
diagonalisation.  CPU only."""
import numpy as np
import pytest

LOG_TGRID = dict(vmin=1e-10, vmax=2.0, n=200)


def _rgflow(oracle, **kw):
    return oracle.run_rgflow(oracle.driver(**kw))


def test_free_chain_first_iteration_closed_form(oracle_mod, golden):
    # examples/freeModel/main.cpp: spinhalf(0,0) impurity coupled by t0 = hopping(0, 2) to the first
    # site: single-particle levels +-t0 per spin -> {+-2t0, +-t0 (x4), 0 (x6)}
    t0 = oracle_mod.hopping(0, 2.0)
    exact = np.sort([-2 * t0, 2 * t0] + [-t0] * 4 + [t0] * 4 + [0.0] * 6)
    eig = _rgflow(oracle_mod, imp_kind=0, bath_kind=0, Lambda=2.0, first_V=0.0, nsites=1, K=1024)[0]
    assert np.abs(eig - exact).max() < 1e-14
    ref = golden("free_chain_L2_K1024")["Eigenvalues0"]
    assert np.abs(ref - exact).max() < 1e-12


@pytest.mark.parametrize("key,imp_kind", [("free_chain_L2_K1024", 0), ("kondo_J025_L2_K1024", 1)])
def test_rgflow_matches_reference_outputs(oracle_mod, golden, key, imp_kind):
    """All 41 iterations: identical state counts (so identical truncation decisions and kept counts)
    and eigenvalues equal to 1e-8 absolute (different LAPACK builds accumulate rounding along the
    chain; the measured difference is ~5e-10)."""
    g = golden(key)
    eigs = _rgflow(oracle_mod, imp_kind=imp_kind, J=0.25, S=0.5, bath_kind=0, Lambda=2.0, first_V=0.0, nsites=41,
                   K=1024)
    for n in range(41):
        ref = g["Eigenvalues%d" % n]
        assert len(ref) == len(eigs[n]), n
        assert np.abs(ref - eigs[n]).max() < 1e-8, n


def _thermo_case(oracle, key):
    if key == "free_chain_L2_K1024":
        return dict(drv=dict(imp_kind=0, bath_kind=0, Lambda=2.0, first_V=0.0, nsites=41, K=1024), min_iter0=1)
    if key == "kondo_J025_L2_K1024":
        return dict(drv=dict(imp_kind=1, J=0.25, S=0.5, bath_kind=0, Lambda=2.0, first_V=0.0, nsites=41, K=1024),
                    min_iter0=1)
    if key == "siam_U02_L25_K512":
        # recovered parameters: eps = -U/2, K = 1024, V = sqrt(2 Gamma / pi) (no A_Lambda factor)
        return dict(drv=dict(imp_kind=0, eps=-0.1, U=0.2, bath_kind=0, Lambda=2.5,
                             first_V=np.sqrt(2 * 0.01 / np.pi), nsites=41, K=1024), min_iter0=0)
    if key == "siam_zero_L25_K512":
        # recovered: the free-chain driver form (no V site, rescale sqrt(Lambda) from site 0), K = 1024
        return dict(drv=dict(imp_kind=0, bath_kind=0, Lambda=2.5, first_V=0.0, nsites=41, K=1024), min_iter0=0)
    raise KeyError(key)


@pytest.mark.parametrize("key", ["free_chain_L2_K1024", "kondo_J025_L2_K1024", "siam_U02_L25_K512",
                                 "siam_zero_L25_K512"])
def test_thermodynamics_matches_reference_outputs(oracle_mod, golden, key):
    """Entropy and specific heat on the reference's 200-point grid.  Below T ~ 3e-7 both codes are
    dominated by the exp(-100) clamp of regulateExp (fdmThermodynamics.hpp:240-249) times the
    d^k shell prefactors; there the shipped files disagree with this code version, so the
    comparison is made where the clamp does not dominate."""
    g = golden(key)
    T = g["temperatureArray"]
    assert np.abs(T - oracle_mod.temperature_grid(**LOG_TGRID)).max() < 1e-15
    case = _thermo_case(oracle_mod, key)
    S, Cv = oracle_mod.run_thermo(oracle_mod.driver(**case["drv"]), T, case["min_iter0"])
    m = T >= 3e-7
    assert np.abs(S - g["entropy"])[m].max() < 1e-8
    rel = np.abs(Cv - g["specificHeat"])[m] / np.maximum(np.abs(g["specificHeat"][m]), 1e-3)
    assert rel.max() < 2e-6


# ---------------------------------------------------------------- brute-force exact diagonalisation
def _jw_ops(n_orb):
    fd = np.array([[0.0, 0.0], [1.0, 0.0]])
    sz = np.diag([1.0, -1.0])
    ops = []
    for i in range(n_orb):
        m = np.eye(1)
        for j in range(n_orb):
            if j < i:
                m = np.kron(m, sz)
            elif j == i:
                m = np.kron(m, fd)
            else:
                m = np.kron(m, np.eye(2))
        ops.append(m)
    return ops


def _ed_siam(eps, U, V, ts, rescales):
    """Spectra of H_0 = eps n_d + U n_up n_dn + V sum_s (d+ c0 + h.c.), then
    H_{k+1} = r_k H_k + t_k sum_s (c_k+ c_{k+1} + h.c.), by dense diagonalisation of the full
    Jordan-Wigner Fock space (a textbook fermion chain, independent of the reference's code)."""
    n_sites = len(ts) + 1
    n_orb = 2 * (n_sites + 1)
    c = _jw_ops(n_orb)  # creation operators
    a = [x.T for x in c]
    nd = [c[s] @ a[s] for s in (0, 1)]
    H = eps * (nd[0] + nd[1]) + U * (nd[0] @ nd[1])
    for s in (0, 1):
        hop = c[s] @ a[2 + s]
        H = H + V * (hop + hop.T)
    specs = []
    active = [0, 1, 2, 3]
    specs.append(np.sort(np.linalg.eigvalsh(H)))  # all orbitals present; extra sites idle (x4 each)
    for k, (t, r) in enumerate(zip(ts, rescales)):
        H = r * H
        for s in (0, 1):
            hop = c[2 + 2 * k + s] @ a[2 + 2 * (k + 1) + s]
            H = H + t * (hop + hop.T)
        specs.append(np.sort(np.linalg.eigvalsh(H)))
    return specs, n_orb


def test_siam_chain_equals_exact_diagonalisation(oracle_mod):
    """Untruncated oracle iterations 0..2 (16, 64, 256 states) against the exact spectrum of the
    Jordan-Wigner chain.  The idle orbitals of the ED space multiply each level by 4 per idle site."""
    eps, U, V, L = -0.25, 0.5, 0.3, 2.5
    ts = [oracle_mod.hopping(0, L), oracle_mod.hopping(1, L)]
    rs = [1.0, np.sqrt(L)]
    specs, n_orb = _ed_siam(eps, U, V, ts, rs)
    imp = oracle_mod.Model("spinhalf", eps=eps, U=U)
    bath = oracle_mod.Model("spinhalf")
    ch = oracle_mod.Chain(imp, bath, 100000)
    ch.add_bath_site([V, V], 1.0)
    ch.update_internal_state()
    got = [ch.all_eigenvalue()]
    for t, r in zip(ts, rs):
        ch.add_bath_site([t, t], r)
        ch.update_internal_state()
        got.append(ch.all_eigenvalue())
    # the ED spectrum carries a 4-fold copy per idle (not yet coupled, zero-energy) site
    for k, g in enumerate(got):
        idle = (n_orb // 2) - (k + 2)
        expect = np.repeat(g, 4 ** idle) if idle > 0 else g
        np.testing.assert_allclose(np.sort(expect), specs[k], atol=1e-12)


def _ed_two_channel_spectra(eps, U, t0, ts, rescales):
    """Spectra of the two-channel Anderson chain by dense diagonalisation of its Jordan-Wigner Fock
    space, built afresh for every length (no idle orbitals).  Orbitals: d_up = 0, d_dn = 1, then four
    per Wilson site, p = 0..3 = (channel 1 up, channel 1 down, channel 2 up, channel 2 down), the
    order of the twoChannel site's four f-dagger operators (spin = p mod 2).
    H_0 = eps n_d + U n_up n_dn + sum_p t0[p] (d+_{p mod 2} c_{0,p} + h.c.);
    H_{k+1} = r_k H_k + t_k sum_p (c+_{k,p} c_{k+1,p} + h.c.)."""
    specs = []
    for n_sites in range(1, len(ts) + 2):
        n_orb = 2 + 4 * n_sites
        c = _jw_ops(n_orb)
        a = [x.T for x in c]
        nd = [c[s] @ a[s] for s in (0, 1)]
        H = eps * (nd[0] + nd[1]) + U * (nd[0] @ nd[1])
        for p in range(4):
            hop = c[p % 2] @ a[2 + p]
            H = H + t0[p] * (hop + hop.T)
        for k in range(n_sites - 1):
            H = rescales[k] * H
            for p in range(4):
                hop = c[2 + 4 * k + p] @ a[2 + 4 * (k + 1) + p]
                H = H + ts[k] * (hop + hop.T)
        specs.append(np.sort(np.linalg.eigvalsh(H)))
    return specs


@pytest.mark.parametrize("t0", [(0.09, 0.09, 0.09, 0.09), (0.11, 0.11, 0.05, 0.05)])
def test_two_channel_chain_equals_exact_diagonalisation(oracle_mod, t0):
    """The two-channel chain (examples/twoChannelSiam/main.cpp: spinhalf impurity with its two f-dagger
    operators duplicated to four, twoChannel bath of 16 states) untruncated at iterations 0 and 1 (64
    and 1024 states) against exact diagonalisation; also with unequal channel couplings
    {V_L, V_L, V_R, V_R} (examples/fdmSpectrumTwoChannel/main.cpp:71), which distinguishes the channels."""
    eps, U, L = -0.2, 0.4, 3.0
    ts = [oracle_mod.hopping(0, L)]
    rs = [1.0]
    specs = _ed_two_channel_spectra(eps, U, t0, ts, rs)
    ch = oracle_mod.Chain(oracle_mod.Model("spinhalf", eps=eps, U=U, dup_ops=True),
                          oracle_mod.Model("two_channel"), 100000)
    ch.add_bath_site(list(t0), 1.0)
    ch.update_internal_state()
    got = [ch.all_eigenvalue()]
    ch.add_bath_site([ts[0]] * 4, rs[0])
    ch.update_internal_state()
    got.append(ch.all_eigenvalue())
    assert [len(g) for g in got] == [64, 1024]
    for k in range(2):
        np.testing.assert_allclose(got[k], specs[k], atol=1e-12)


def _ed_siam_final(eps, U, V, ts, rescales):
    """Final Hamiltonian of the chain of _ed_siam and the impurity creation operators d+_up, d+_dn."""
    n_orb = 2 * (len(ts) + 2)
    c = _jw_ops(n_orb)
    a = [x.T for x in c]
    nd = [c[s] @ a[s] for s in (0, 1)]
    H = eps * (nd[0] + nd[1]) + U * (nd[0] @ nd[1])
    for s in (0, 1):
        hop = c[s] @ a[2 + s]
        H = H + V * (hop + hop.T)
    for k, (t, r) in enumerate(zip(ts, rescales)):
        H = r * H
        for s in (0, 1):
            hop = c[2 + 2 * k + s] @ a[2 + 2 * (k + 1) + s]
            H = H + t * (hop + hop.T)
    return H, c[0], c[1]


def test_fdm_last_shell_equals_exact_diagonalisation(oracle_mod):
    """With K equal to the 256 states of iteration 1, iteration 2 is the first (and last) one that
    checkHigherEnergyDiscarded (nrgcore.hpp:384-386) reports, yet nothing has been discarded: it
    stores one FDM shell over the full, exact eigenbasis (1024 states), whose density matrix is the
    ground-state projector (fdmSpectrum.hpp:122-155, 165-199).
    Its binned weights (fdmSpectrum.hpp:207-277) are then the Lehmann weights of the exact chain:
    |<n|d+|g>|^2 at E_n - E_g (positive) and |<n|d|g>|^2 at E_g - E_n (negative), averaged over the
    degenerate ground states, in bins int(ln(1 + |s dE| / 1e-10) delE), s = 1/Lambda; and each
    spin's total weight is <d d+ + d+ d> = 1."""
    eps, U, G, L, nb = -0.2, 0.5, 0.02, 2.0, 400
    V = np.sqrt(2 * G / np.pi)
    drv = oracle_mod.driver(imp_kind=0, eps=eps, U=U, bath_kind=0, Lambda=L, first_V=V, nsites=3, K=256)
    pos, neg, tr, sums = oracle_mod.run_fdm(drv, nb)
    assert len(tr) == 1
    np.testing.assert_allclose(tr, 1.0, atol=1e-13)
    ts = [oracle_mod.hopping(n, L) for n in range(3)]
    H, cu, cd = _ed_siam_final(eps, U, V, ts, [1.0, np.sqrt(L), np.sqrt(L)])
    w, Q = np.linalg.eigh(H)
    gs = np.nonzero(np.abs(w - w[0]) < 1e-5)[0]
    scale = 1.0 / L
    dele = (nb - 1) / np.log(10 / 1e-10)
    for p, cdag in enumerate((cu, cd)):
        M = Q.T @ cdag @ Q  # <n|d+|m>
        ep, en = np.zeros(nb), np.zeros(nb)
        for g in gs:
            for n in range(len(w)):
                x = np.log(1 + abs(scale * (w[n] - w[g])) / 1e-10) * dele
                assert abs(x - round(x)) > 1e-6 or round(x) == 0  # no weight on a bin edge
                k = int(x)
                if k < nb:
                    ep[k] += M[n, g] ** 2 / len(gs)
                    en[k] += M[g, n] ** 2 / len(gs)
        np.testing.assert_allclose(pos[p].sum() + neg[p].sum(), 1.0, atol=1e-12)
        np.testing.assert_allclose(pos[p], ep, atol=1e-12)
        np.testing.assert_allclose(neg[p], en, atol=1e-12)


def test_truncation_keeps_K_plus_one_and_shifts_only_when_truncating(oracle_mod):
    imp = oracle_mod.Model("spinhalf", eps=-0.25, U=0.5)
    bath = oracle_mod.Model("spinhalf")
    ch = oracle_mod.Chain(imp, bath, 40)
    ch.add_bath_site([0.2, 0.2], 1.0)
    ch.update_internal_state()            # 16 states <= 40: no truncation, no shift
    assert ch.kept_counts().sum() == 16 and len(ch.relative_gs()) == 0
    ch.add_bath_site([oracle_mod.hopping(0, 2.5)] * 2, 1.0)
    ch.update_internal_state()            # 64 states > 40: keep E < sorted[40] + 1e-8
    allE = ch.all_eigenvalue()
    kept = ch.kept_counts().sum()
    assert kept >= 41
    assert kept == np.sum(allE < allE[40] + 1e-8)
    assert len(ch.relative_gs()) == 1 and ch.relative_gs()[0] == allE[0]
    # the shifted spectrum starts at 0
    assert min(e.min() for e in ch.eigenvalues_q() if len(e)) == 0.0


def test_fdm_reduced_density_matrices_have_unit_trace(oracle_mod):
    """The backward sweep hands each shell the partial trace of the previous shell's density matrix
    over the bath site (fdmSpectrum.hpp:71-114) plus, at the last shell, the ground-state projector:
    every shell's rhoZero then has trace 1 (fdmSpectrum.hpp:191-196 prints it)."""
    drv = oracle_mod.driver(imp_kind=0, eps=-0.15, U=0.3, bath_kind=0, Lambda=2.0,
                            first_V=np.sqrt(2 * 0.01 / np.pi), nsites=14, K=96)
    pos, neg, tr, sums = oracle_mod.run_fdm(drv, 200)
    assert len(tr) >= 8
    np.testing.assert_allclose(tr, 1.0, atol=1e-12)
    assert np.all(pos >= 0) and np.all(neg >= 0)


@pytest.mark.parametrize("U,L,K,nsites", [(0.3, 2.0, 256, 18), (0.6, 2.5, 256, 14), (0.3, 2.0, 1024, 12)])
def test_fdm_complete_basis_sum_rule(oracle_mod, U, L, K, nsites):
    """The multi-shell backward pass against the complete-basis sum rule: the signed weights
    w+ + w- of rhoDotOperators (fdmSpectrum.hpp:231-243), summed over every shell before the |.|, the
    1e-20 cut and the bin-range test, add up to <{B, B+}> = 1 for B = d+_sigma chi.  Every discarded
    state of every recorded shell and every state of the last shell enter once, so a wrong density
    matrix rotation (U^T rho U instead of U rho U^T, fdmSpectrum.hpp:87-91) or a term dropped from the
    partial trace (:95-113) breaks it, while the trace of each shell's rho would still be 1 after the
    first.  K is a power of the bath dimension 4: the reference records a shell only when the previous
    iteration kept at least K states (checkHigherEnergyDiscarded, nrgcore.hpp:384-386), so otherwise
    the first truncating iteration is never recorded and its discarded weight is missing from the sum
    (test_fdm_first_truncating_shell_is_not_recorded)."""
    drv = oracle_mod.driver(imp_kind=0, eps=-U / 2, U=U, bath_kind=0, Lambda=L,
                            first_V=np.sqrt(2 * 0.01 / np.pi), nsites=nsites, K=K)
    _, _, tr, _, sg = oracle_mod.run_fdm(drv, 200, signed=True)
    assert len(tr) >= 6
    np.testing.assert_allclose(sg, 1.0, atol=1e-12)
    try:
        for plant in (1, 2):
            oracle_mod.set_fdm_plant(plant)
            _, _, _, _, bad = oracle_mod.run_fdm(drv, 200, signed=True)
            assert np.abs(bad - 1.0).max() > 1e-3, plant
    finally:
        oracle_mod.set_fdm_plant(0)


def test_fdm_first_truncating_shell_is_not_recorded(oracle_mod):
    """With K = 96 the first iteration that truncates (256 states) follows one that kept 64 < K, so
    checkHigherEnergyDiscarded is false there and the shell is not recorded (nrgcore.hpp:384-386,
    examples/fdmSpectrumSiam/main.cpp:130-140): the signed total falls short of 1 by the same amount
    whatever the chain length, the weight of the discarded states of that one iteration."""
    short = []
    for nsites in (10, 14, 20):
        drv = oracle_mod.driver(imp_kind=0, eps=-0.15, U=0.3, bath_kind=0, Lambda=2.0,
                                first_V=np.sqrt(2 * 0.01 / np.pi), nsites=nsites, K=96)
        short.append(1.0 - oracle_mod.run_fdm(drv, 200, signed=True)[4][0])
    assert 1e-5 < short[0] < 1e-2
    assert abs(short[2] - short[1]) < 0.02 * short[0]


# ---------------------------------------------------------------- static observables
def test_thermal_expectation_equals_exact_diagonalisation(oracle_mod):
    """With K equal to the 256 states of iteration 1, the one recorded shell is the full exact
    eigenbasis of the 10-orbital chain, with every state discarded and w_m = 1.  Then
    qsysOPeratorExpValue (fdmThermodynamics.hpp:191-206) is the thermal average
    sum_l x_l <l|O|l> / sum_l x_l over the exact Jordan-Wigner chain, x_l = exp((E_0 - E_l) s / T) with
    s = 1/Lambda and the exponent clamped at -100 (regulateExp, fdmThermodynamics.hpp:240-249), and
    the entropy is ln sum_l x_l + <(E - E_0) s>/T.  Operators: n_up n_dn, n_up, n_dn and the identity."""
    eps, U, G, L = -0.2, 0.5, 0.02, 2.0
    V = np.sqrt(2 * G / np.pi)
    drv = oracle_mod.driver(imp_kind=0, eps=eps, U=U, bath_kind=0, Lambda=L, first_V=V, nsites=3, K=256)
    T = np.logspace(-4, 1, 31)
    S, _, occ = oracle_mod.run_thermo_ops(drv, T, 0, 1)
    _, _, dbl = oracle_mod.run_thermo_ops(drv, T, 0, 0)
    ts = [oracle_mod.hopping(n, L) for n in range(3)]
    H, cu, cd = _ed_siam_final(eps, U, V, ts, [1.0, np.sqrt(L), np.sqrt(L)])
    w, Q = np.linalg.eigh(H)
    nu, nd = cu @ cu.T, cd @ cd.T
    x = np.exp(np.clip((w[0] - w)[None, :] / L / T[:, None], -100, 100))
    Z = x.sum(1)
    for got, O in ((dbl[0], nu @ nd), (occ[0], nu), (occ[1], nd), (occ[2], np.eye(len(w)))):
        diag = np.einsum("ij,ik,kj->j", Q, O, Q)
        np.testing.assert_allclose(got, x @ diag / Z, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(S, np.log(Z) + (x @ ((w - w[0]) / L)) / Z / T, rtol=1e-9, atol=1e-12)


def test_thermal_occupation_at_particle_hole_symmetry(oracle_mod):
    """At eps = -U/2 the truncated chain is particle-hole symmetric and every shell's discarded set is
    closed under the symmetry, so <n_up> = <n_dn> = 1/2 at every temperature; the identity averages
    to 1 (normalised shell weights w_m over Z'_m).  Double occupancy lies in [0, 1/4] for U > 0."""
    U, L = 0.5, 2.5
    drv = oracle_mod.driver(imp_kind=0, eps=-U / 2, U=U, bath_kind=0, Lambda=L,
                            first_V=np.sqrt(2 * 0.02 / np.pi), nsites=20, K=128)
    T = np.logspace(-9, 0, 40)
    _, _, occ = oracle_mod.run_thermo_ops(drv, T, 0, 1)
    _, _, dbl = oracle_mod.run_thermo_ops(drv, T, 0, 0)
    np.testing.assert_allclose(occ[0], 0.5, atol=1e-9)
    np.testing.assert_allclose(occ[1], 0.5, atol=1e-9)
    np.testing.assert_allclose(occ[2], 1.0, atol=1e-12)
    assert np.all(dbl[0] > 0) and np.all(dbl[0] < 0.25)


def test_oracle_timing_with_single_core_tail(oracle_mod):
    """bench.py's CPU baseline: sites timed at the current thread count, then the next sites of the
    same chain at one thread; both spans are positive and the thread count is the caller's to restore."""
    drv = oracle_mod.driver(imp_kind=0, eps=-0.15, U=0.3, bath_kind=0, Lambda=2.0,
                            first_V=float(np.sqrt(2 * 0.01 / np.pi)), nsites=8, K=64)
    s, s1 = oracle_mod.time_forward_with_single_core(drv, 2, 3, 2)
    assert s > 0 and s1 > 0
    s0, s10 = oracle_mod.time_forward_with_single_core(drv, 2, 3, 0)
    assert s0 > 0 and s10 == 0
