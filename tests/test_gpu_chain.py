"""GPU parity: libnrgk (HIP, through its C ABI) against the FP64 oracle and the reference's outputs.

Tolerances (stated per result):
  * sector labels, block dimensions, kept counts: bit-exact;
  * assembled sector Hamiltonians fed the oracle's state (lockstep): bit-exact;
  * eigenvalues of each step vs LAPACK on the same Hamiltonian: <= 1e-10 (iteration energy units);
  * eigenvectors: ||U^T U - I||_max and ||H U - U diag(E)||_max <= 1e-10;
  * chain eigenvalues vs the oracle chain / the reference's files: <= 1e-8 absolute;
  * thermodynamics vs the reference's files: as for the oracle (where the exp clamp does not dominate);
  * FDM weights: total weight to 1e-9 relative, broadened spectrum to 1e-6 of its maximum.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state_from_oracle(nk, o):
    """The previous-basis state an assembly reads, exported from the oracle chain (after
    update_internal_state): sectors, kept energies, kept x kept f-dagger blocks."""
    qn = o.pre_qn()
    kept = o.kept_counts()
    eq = o.eigenvalues_q()
    energies = np.concatenate([eq[i][:kept[i]] for i in range(len(kept))]) if kept.sum() else np.zeros(0)
    ops = []
    for blocks in o.fdag():
        ops.append([(i, j, np.ascontiguousarray(m[:kept[i], :kept[j]])) for i, j, m in blocks])
    return nk.Site.from_arrays(qn, kept, energies, np.ones(len(kept)), ops)


def _models(nk, oracle, kind, eps, U):
    if kind == "siam":
        return (nk.Site.spinhalf(eps, U), nk.Site.spinhalf(0, 0),
                oracle.Model("spinhalf", eps=eps, U=U), oracle.Model("spinhalf"), 2)
    return (nk.Site.spinhalf(eps, U).duplicate_ops(), nk.Site.two_channel(),
            oracle.Model("spinhalf", eps=eps, U=U, dup_ops=True), oracle.Model("two_channel"), 4)


def _check_eigen_step(g, tol=1e-10):
    qn, dims = g.sectors()
    eq = g.eigenvaluesQ
    worst = 0.0
    for s, n in enumerate(dims):
        if n == 0:
            continue
        H = g.block_h(s)
        w = np.linalg.eigvalsh(H)
        assert np.abs(w - eq[s]).max() <= tol, (s, n, np.abs(w - eq[s]).max())
        U = g.current_hamiltonQ(s)
        assert np.abs(U.T @ U - np.eye(n)).max() <= tol
        assert np.abs(H @ U - U * eq[s]).max() <= tol
        worst = max(worst, np.abs(w - eq[s]).max())
    return worst


@pytest.mark.parametrize("kind,K,nit", [("siam", 200, 10), ("two_channel", 200, 3)])
def test_lockstep_assembly_bit_exact_and_eigen_accuracy(nk, oracle_mod, kind, K, nit):
    imp, bath, oimp, obath, nops = _models(nk, oracle_mod, kind, -0.2, 0.4)
    g = nk.NrgCore(imp, bath, K)
    g.keep_hamiltonians(True)
    o = oracle_mod.Chain(oimp, obath, K, keep_h=True)
    L = 3.0 if kind == "two_channel" else 2.0
    V = np.sqrt(2 * 0.01 / np.pi)
    ts = [V] + [oracle_mod.hopping(n, L) for n in range(nit)]
    rs = [1.0] + [1.0 if n == 0 else np.sqrt(L) for n in range(nit)]
    for it, (t, r) in enumerate(zip(ts, rs)):
        if it > 0:
            g.set_state(_state_from_oracle(nk, o), o.cnt)
        g.add_bath_site([t] * nops, r)
        o.add_bath_site([t] * nops, r)
        qn, dims = g.sectors()
        assert np.array_equal(qn, o.cur_qn()) and np.array_equal(dims, o.sector_dims())
        for s, n in enumerate(dims):
            if n:
                hg, ho = g.block_h(s), o.block_h(s, n)
                assert np.array_equal(hg, ho), (it, s, np.abs(hg - ho).max())
        _check_eigen_step(g)
        g.update_internal_state()
        o.update_internal_state()
        assert np.array_equal(g.kept_counts(), o.kept_counts()), it


def test_free_running_chain_matches_oracle(nk, oracle_mod):
    """config 1 form (SIAM U=0.5, eps=-0.25, Gamma=0.02, Lambda=2.5) at K=256, 14 sites."""
    K, L = 256, 2.5
    V = np.sqrt(2 * 0.02 / np.pi)
    g = nk.NrgCore(nk.Site.spinhalf(-0.25, 0.5), nk.Site.spinhalf(0, 0), K)
    g.keep_hamiltonians(True)
    o = oracle_mod.Chain(oracle_mod.Model("spinhalf", eps=-0.25, U=0.5), oracle_mod.Model("spinhalf"), K)
    ts = [V] + [oracle_mod.hopping(n, L) for n in range(14)]
    rs = [1.0] + [1.0 if n == 0 else np.sqrt(L) for n in range(14)]
    for it, (t, r) in enumerate(zip(ts, rs)):
        g.add_bath_site([t, t], r)
        o.add_bath_site([t, t], r)
        _check_eigen_step(g)
        g.update_internal_state()
        o.update_internal_state()
        assert np.array_equal(g.kept_counts(), o.kept_counts()), it
        a, b = g.all_eigenvalue, o.all_eigenvalue()
        assert len(a) == len(b) and np.abs(a - b).max() < 1e-8, (it, np.abs(a - b).max())
        assert np.array_equal(g.relativeGroundStateEnergy.shape, o.relative_gs().shape)
        # Wilson-site f-dagger: kept x kept blocks are gauge-dependent, their Frobenius norms are not
        fg, fo = g.get_f_dag_operator(), o.fdag()
        kept = o.kept_counts()
        for bg, bo in zip(fg, fo):
            assert [(i, j) for i, j, _ in bg] == [(i, j) for i, j, _ in bo]
            for (_, _, mg), (i, j, mo) in zip(bg, bo):
                ng, no = np.linalg.norm(mg), np.linalg.norm(mo[:kept[i], :kept[j]])
                assert abs(ng - no) <= 1e-9 * max(1.0, no)


@pytest.mark.parametrize("key,kondo", [("free_chain_L2_K1024", False), ("kondo_J025_L2_K1024", True)])
def test_reference_rgflow_outputs_on_gpu(nk, oracle_mod, golden, key, kondo):
    gd = golden(key)
    imp = nk.Site.kondo(0.25, 0.5) if kondo else nk.Site.spinhalf(0, 0)
    g = nk.NrgCore(imp, nk.Site.spinhalf(0, 0), 1024)
    for n in range(41):
        t = oracle_mod.hopping(n, 2.0)
        g.add_bath_site([t, t], np.sqrt(2.0))
        g.update_internal_state()
        ref = gd["Eigenvalues%d" % n]
        a = g.all_eigenvalue
        assert len(a) == len(ref), n
        assert np.abs(a - ref).max() < 1e-8, (n, np.abs(a - ref).max())


def test_reference_thermodynamics_on_gpu(nk, golden):
    """examples/entropySIAM driver (recovered parameters of resultSIAMU0.200000.h5)."""
    gd = golden("siam_U02_L25_K512")
    T = gd["temperatureArray"]
    L = 2.5
    g = nk.NrgCore(nk.Site.spinhalf(-0.1, 0.2), nk.Site.spinhalf(0, 0), 1024)
    rec = nk.FdmRecorder(g, None)
    V = np.sqrt(2 * 0.01 / np.pi)
    g.add_bath_site([V, V], 1.0)
    g.update_internal_state()
    min_it = 0
    for n in range(41):
        t = nk.hopping(n, L)
        g.add_bath_site([t, t], 1.0 if n == 0 else np.sqrt(L))
        if g.checkHigherEnergyDiscarded():
            rec.record()
        else:
            min_it += 1
        g.update_internal_state()
    S, Cv = rec.thermodynamics(L, T, scale_offset=-1, first_cnt=min_it + 1)
    m = T >= 3e-7
    assert np.abs(S - gd["entropy"])[m].max() < 1e-8
    rel = np.abs(Cv - gd["specificHeat"])[m] / np.maximum(np.abs(gd["specificHeat"][m]), 1e-3)
    assert rel.max() < 2e-6


def _gpu_fdm(nk, eps, U, G, L, K, nsites, nbins, signed=False):
    g = nk.NrgCore(nk.Site.spinhalf(eps, U), nk.Site.spinhalf(0, 0), K)
    op = nk.SystemOperator(g, nk.Site.fdag_chi_spinhalf())
    rec = nk.FdmRecorder(g, op, nbins, 1e-10, 10.0)
    V = np.sqrt(2 * G / np.pi)
    g.add_bath_site([V, V], 1.0)
    op.update()
    g.update_internal_state()
    min_it = 0
    for n in range(nsites):
        t = nk.hopping(n, L)
        g.add_bath_site([t, t], 1.0 if n == 0 else np.sqrt(L))
        op.update()
        if g.checkHigherEnergyDiscarded():
            rec.record()
        else:
            min_it += 1
        g.update_internal_state()
    out = rec.backward(L, scale_offset=-1, first_cnt=min_it + 1), rec.energy_points()
    return out + (rec.signed_weights(),) if signed else out


def test_fdm_spectrum_matches_oracle(nk, oracle_mod):
    eps, U, G, L, K, ns, nb = -0.15, 0.3, 0.01, 2.0, 128, 18, 400
    (pos, neg, tr, sums), x = _gpu_fdm(nk, eps, U, G, L, K, ns, nb)
    drv = oracle_mod.driver(imp_kind=0, eps=eps, U=U, bath_kind=0, Lambda=L, first_V=np.sqrt(2 * G / np.pi),
                            nsites=ns, K=K)
    opos, oneg, otr, osums = oracle_mod.run_fdm(drv, nb)
    assert len(tr) == len(otr)
    np.testing.assert_allclose(tr, 1.0, atol=1e-10)  # rhoTrace per shell
    np.testing.assert_allclose(sums, osums, rtol=1e-8)
    tot, otot = pos.sum() + neg.sum(), opos.sum() + oneg.sum()
    assert abs(tot - otot) <= 1e-9 * otot
    # Per spin the weight is near the <{d, d+}> = 1 sum rule but not on it: the reference bins
    # |(B rho)_vv' A_v'v| element by element (fdmSpectrum.hpp:237-258), so contributions of the
    # off-diagonal reduced density matrix add with either sign, a few per cent over one here.
    for p in range(2):
        assert 0.9 < pos[p].sum() + neg[p].sum() < 1.1
    # broadened spectra (gauge- and bin-edge-robust)
    om = np.logspace(-8, 0, 200)
    for p in range(2):
        for a, b in ((pos[p], opos[p]), (neg[p], oneg[p])):
            ag, ao = nk.broaden(x, a, om, 0.6), oracle_mod.broaden(x, b, om, 0.6)
            assert np.abs(ag - ao).max() <= 1e-6 * max(1e-12, np.abs(ao).max())


def _occupation_ops(nk):
    """n_up, n_dn and the identity on the spinhalf impurity's sectors (1x1 blocks, zero blocks left
    out as the reference's get_block_operators does, fermionBasis.hpp:291-318)."""
    qn, dims, en, chi, _ = nk.Site.spinhalf(0.0, 0.0).export()
    assert np.all(dims == 1)
    ops = [[(i, i, np.array([[float(qn[i][0])]])) for i in range(len(dims)) if qn[i][0]],
           [(i, i, np.array([[float(qn[i][1])]])) for i in range(len(dims)) if qn[i][1]],
           [(i, i, np.eye(1)) for i in range(len(dims))]]
    return nk.Site.from_arrays(qn, dims, en, chi, ops)


def test_thermal_expectation_values_match_oracle(nk, oracle_mod):
    """fdmThermodynamics with system operators (setSystemOperator / qsysOPeratorExpValue,
    fdmThermodynamics.hpp:39-41, 63-87, 191-211) on a truncated SIAM chain: double occupancy,
    n_up, n_dn and the identity against the oracle to 1e-8 absolute (eigenvalues to 1e-10; sums over
    degenerate multiplets are basis-independent), the identity to 1 within 1e-12."""
    eps, U, G, L, K, ns = -0.15, 0.3, 0.01, 2.0, 128, 24
    V = np.sqrt(2 * G / np.pi)
    T = np.logspace(-10, 0, 41)
    g = nk.NrgCore(nk.Site.spinhalf(eps, U), nk.Site.spinhalf(0, 0), K)
    dbl = nk.SystemOperator(g, nk.Site.double_occupancy_spinhalf())
    occ = nk.SystemOperator(g, _occupation_ops(nk))
    rec_d, rec_o = nk.FdmRecorder(g, None), nk.FdmRecorder(g, None)
    rec_d.setSystemOperator(dbl)
    rec_o.setSystemOperator(occ)
    g.add_bath_site([V, V], 1.0)
    dbl.update()
    occ.update()
    g.update_internal_state()
    min_it = 0
    for n in range(ns):
        t = nk.hopping(n, L)
        g.add_bath_site([t, t], 1.0 if n == 0 else np.sqrt(L))
        dbl.update()
        occ.update()
        if g.checkHigherEnergyDiscarded():
            rec_d.record()
            rec_o.record()
        else:
            min_it += 1
        g.update_internal_state()
    vd = rec_d.expectation_values(L, T, scale_offset=-1, first_cnt=min_it + 1)
    vo = rec_o.expectation_values(L, T, scale_offset=-1, first_cnt=min_it + 1)
    S, Cv = rec_o.thermodynamics(L, T, scale_offset=-1, first_cnt=min_it + 1)
    drv = oracle_mod.driver(imp_kind=0, eps=eps, U=U, bath_kind=0, Lambda=L, first_V=V, nsites=ns, K=K)
    oS, oC, od = oracle_mod.run_thermo_ops(drv, T, 0, 0)
    _, _, oo = oracle_mod.run_thermo_ops(drv, T, 0, 1)
    np.testing.assert_allclose(vo[2], 1.0, atol=1e-12)
    assert np.abs(vd - od).max() < 1e-8, np.abs(vd - od).max()
    assert np.abs(vo - oo).max() < 1e-8, np.abs(vo - oo).max()
    assert np.abs(S - oS).max() < 1e-8
    # the recorder without a system operator refuses the expectation values
    with pytest.raises(RuntimeError):
        nk.FdmRecorder(g, None).expectation_values(L, T)


def test_z_averaged_thermodynamics_matches_oracle(nk, oracle_mod):
    """z-averaging (the z-averaged SIAM workload at a smaller K and length: eps = -0.25, U = 0.5,
    Gamma = 0.015, Lambda = 4): one chain per twist z = k/4 with the hoppings of the z-shifted
    discretisation (times A_Lambda, so that z = 1 is Wilson's chain), thermodynamics per chain
    against the oracle driven with the same hoppings (entropy to 1e-8 absolute where the exp clamp
    does not dominate), then averaged over z."""
    L, K, ns = 4.0, 256, 18
    eps, U, G = -0.25, 0.5, 0.015
    V = np.sqrt(2 * G / np.pi)
    A = 0.5 * np.log(L) * (1 + 1 / L) / (1 - 1 / L)
    T = np.logspace(-6, 0, 25)
    S_avg = np.zeros(len(T))
    for z in (0.25, 0.5, 0.75, 1.0):
        hop = nk.z_chain(z, L, ns)[0] * A
        g = nk.NrgCore(nk.Site.spinhalf(eps, U), nk.Site.spinhalf(0, 0), K)
        rec = nk.FdmRecorder(g, None)
        g.add_bath_site([V, V], 1.0)
        g.update_internal_state()
        min_it = 0
        for n in range(ns):
            g.add_bath_site([hop[n], hop[n]], 1.0 if n == 0 else np.sqrt(L))
            if g.checkHigherEnergyDiscarded():
                rec.record()
            else:
                min_it += 1
            g.update_internal_state()
        S, Cv = rec.thermodynamics(L, T, scale_offset=-1, first_cnt=min_it + 1)
        drv = oracle_mod.driver(imp_kind=0, eps=eps, U=U, bath_kind=0, Lambda=L, first_V=V, nsites=ns, K=K)
        oS, oC = oracle_mod.run_thermo_hoppings(drv, hop, T, 0)
        m = T >= 1e-5
        assert np.abs(S - oS)[m].max() < 1e-8, (z, np.abs(S - oS)[m].max())
        S_avg += S / 4
    assert np.all(np.isfinite(S_avg))


def _jw(n_orb):
    D = 1 << n_orb
    ops = []
    for i in range(n_orb):
        m = np.zeros((D, D))
        for st in range(D):
            if not (st >> i) & 1:
                m[st | (1 << i), st] = (-1) ** bin(st & ((1 << i) - 1)).count("1")
        ops.append(m)
    return ops


def _chain_vs_ed(nk, imp, bath, H_imp, links, n_imp, n_sites, L):
    """Untruncated chain of a given impurity and two-orbital bath against exact diagonalisation of
    the whole fermion chain: H_0 = H_imp, H_{k+1} = r_k H_k + t_k sum_c (a+_c b_c + h.c.), where a_c
    are the impurity (k = 0) or previous site's orbitals of channel c and b_c the new site's."""
    g = nk.NrgCore(imp, bath, 100000)
    n_orb = n_imp + 2 * n_sites
    f = _jw(n_orb)
    H = np.zeros((1 << n_orb, 1 << n_orb))
    H[...] = H_imp(f)
    prev = list(links)
    for k in range(n_sites):
        t, r = nk.hopping(k, L), (1.0 if k == 0 else np.sqrt(L))
        g.add_bath_site([t, t], r)
        new = [n_imp + 2 * k, n_imp + 2 * k + 1]
        H = r * H
        for a, b in zip(prev, new):
            hop = f[a] @ f[b].T
            H = H + t * (hop + hop.T)
        prev = new
        g.update_internal_state()  # all_eigenvalue is refreshed there (nrgcore.hpp:105-109, 269-327)
        got = np.sort(g.all_eigenvalue)
        # the ED space holds the sites not yet added as idle orbitals: 4 copies per idle site
        ed = np.linalg.eigvalsh(H)
        idle = n_sites - 1 - k
        exp = np.sort(np.repeat(got, 4 ** idle))
        assert np.abs(exp - ed).max() < 1e-9, (k, np.abs(exp - ed).max())


def test_resonant_level_chain_equals_exact_diagonalisation(nk):
    """examples/resonantLevel/main.cpp with the hopping list sized to the two lead operators (the
    driver passes four and would throw): resonantTwoLead impurity, spinnlessTwoLead bath, Lambda = 3."""
    eps, vl, vr, U = -0.025, 0.11, 0.07, 0.05

    def H_imp(f):
        n = [x @ x.T for x in f[:3]]
        I = np.eye(f[0].shape[0])
        return eps * n[0] + vl * (f[0] @ f[1].T + f[1] @ f[0].T) + vr * (f[0] @ f[2].T + f[2] @ f[0].T) \
            + U * (n[0] - I) @ (n[1] + n[2] - I)

    _chain_vs_ed(nk, nk.Site.resonant_two_lead(eps, vl, vr, U), nk.Site.spinless_two_lead(), H_imp, [1, 2], 3, 3, 3.0)


def test_superconducting_impurity_chain_equals_exact_diagonalisation(nk):
    """spinhalfSC with a pairing term on a spin-only free chain (only S_z conserved)."""
    e, u, h, d = -0.1, 0.3, 0.02, 0.05

    def H_imp(f):
        nu, nd = f[0] @ f[0].T, f[1] @ f[1].T
        return e * (nu + nd) + h * (nu - nd) + u * nu @ nd + d * (f[0].T @ f[1].T + f[1] @ f[0])

    _chain_vs_ed(nk, nk.Site.spinhalf_sc(e, u, h, d), nk.Site.free("spin_only"), H_imp, [0, 1], 2, 3, 2.0)


def test_fdm_recorder_saves_the_reference_layout(nk, tmp_path):
    """NrgData::saveCurrentData + saveqOperator (nrgData.hpp:134-205) from the recorder: every dataset
    read back equals what the chain and the observable held when the shell was recorded."""
    L, K = 2.0, 64
    g = nk.NrgCore(nk.Site.spinhalf(-0.15, 0.3), nk.Site.spinhalf(0, 0), K)
    op = nk.SystemOperator(g, nk.Site.fdag_chi_spinhalf())
    rec = nk.FdmRecorder(g, op, 200, 1e-10, 10.0)
    V = np.sqrt(2 * 0.01 / np.pi)
    g.add_bath_site([V, V], 1.0)
    op.update()
    g.update_internal_state()
    seen = {}
    for n in range(7):
        t = nk.hopping(n, L)
        prev_qn, prev_kept = g.current_sysmQ, g.kept_counts()
        g.add_bath_site([t, t], 1.0 if n == 0 else np.sqrt(L))
        op.update()
        if g.checkHigherEnergyDiscarded():
            rec.record()
            qn, dims = g.sectors()
            seen[g.nrg_iterations_cnt] = (qn, dims, g.eigenvaluesQ, [g.current_hamiltonQ(s) for s in range(len(dims))],
                                          prev_qn, prev_kept, op.blocks())
        g.update_internal_state()
    assert len(seen) >= 3
    path = str(tmp_path / "nrgdata.h5")
    rec.save(path)
    for cnt, (qn, dims, eq, U, pqn, pk, ops) in seen.items():
        gr = "/NrgItr%d/" % cnt
        for s in range(len(dims)):
            np.testing.assert_array_equal(nk.h5_read(path, gr + "current_sysmQ%d" % s), qn[s])
            np.testing.assert_array_equal(nk.h5_read(path, gr + "eigenvaluesQ%d" % s), eq[s])
            np.testing.assert_array_equal(nk.h5_read(path, gr + "current_hamiltonQ%d" % s), U[s].reshape(-1))
        for s in range(len(pqn)):
            np.testing.assert_array_equal(nk.h5_read(path, gr + "pre_sysmQ%d" % s), pqn[s])
            np.testing.assert_array_equal(nk.h5_read(path, gr + "eigenvaluesQ_kept_indices%d" % s), np.arange(pk[s]))
        allE = np.sort(np.concatenate(eq))
        np.testing.assert_array_equal(nk.h5_read(path, "/Eigenvalues%d" % cnt), allE)
        og = "/nparticle%d/" % (cnt - 1)
        for p, blocks in enumerate(ops):
            for k, (i, j, b) in enumerate(blocks):
                np.testing.assert_array_equal(nk.h5_read(path, og + "at%d/idVector%d" % (p, k)), [i, j])
                np.testing.assert_array_equal(nk.h5_read(path, og + "at%d/qmat%d" % (p, k)), b.reshape(-1))
            np.testing.assert_array_equal(nk.h5_read(path, og + "at%d/rowVector" % p), [b.shape[0] for _, _, b in blocks])


def test_all_eigenvalue_keeps_the_previous_list_until_update(nk):
    """all_eigenvalue is set by update_internal_state (discard_higher_energies, nrgcore.hpp:269-327):
    after the next add_bath_site it still holds that list, whether or not it was read in between."""
    def chain():
        g = nk.NrgCore(nk.Site.spinhalf(-0.15, 0.3), nk.Site.spinhalf(0, 0), 64)
        g.add_bath_site([0.08, 0.08], 1.0)
        g.update_internal_state()
        for n in range(4):
            g.add_bath_site([nk.hopping(n, 2.0)] * 2, 1.0 if n == 0 else np.sqrt(2.0))
            g.update_internal_state()
        return g
    a, b = chain(), chain()
    before = a.all_eigenvalue  # read after update_internal_state
    for g in (a, b):
        g.add_bath_site([nk.hopping(4, 2.0)] * 2, np.sqrt(2.0))
    np.testing.assert_array_equal(a.all_eigenvalue, before)
    np.testing.assert_array_equal(b.all_eigenvalue, before)  # not read before add_bath_site
    b.update_internal_state()
    assert len(b.all_eigenvalue) != len(before) or np.any(b.all_eigenvalue != before)


def test_large_block_crossover_matches_oracle(nk, oracle_mod, monkeypatch):
    """Blocks wider than the batched solver's limit go to the FP64 tridiagonalising solver on a side
    stream while the rest take the batched path: with the limit lowered to 40 rows both kinds occur
    in every iteration; eigenpairs to 1e-10, chain eigenvalues to 1e-8 of the oracle."""
    monkeypatch.setenv("NRGK_MIXED_MAX", "40")
    K, L = 256, 2.5
    V = np.sqrt(2 * 0.02 / np.pi)
    g = nk.NrgCore(nk.Site.spinhalf(-0.25, 0.5), nk.Site.spinhalf(0, 0), K)
    g.keep_hamiltonians(True)
    o = oracle_mod.Chain(oracle_mod.Model("spinhalf", eps=-0.25, U=0.5), oracle_mod.Model("spinhalf"), K)
    ts = [V] + [oracle_mod.hopping(n, L) for n in range(8)]
    rs = [1.0] + [1.0 if n == 0 else np.sqrt(L) for n in range(8)]
    both = 0
    for it, (t, r) in enumerate(zip(ts, rs)):
        g.add_bath_site([t, t], r)
        o.add_bath_site([t, t], r)
        _, dims = g.sectors()
        both += int(dims.max() > 40 and (dims[dims > 0].min() <= 40))
        _check_eigen_step(g)
        g.update_internal_state()
        o.update_internal_state()
        assert np.array_equal(g.kept_counts(), o.kept_counts()), it
        a, b = g.all_eigenvalue, o.all_eigenvalue()
        assert len(a) == len(b) and np.abs(a - b).max() < 1e-8, (it, np.abs(a - b).max())
    assert both >= 4


def test_all_blocks_on_side_streams_match_oracle(nk, oracle_mod, monkeypatch):
    """With the limit at 0 every block is solved on the side streams and the batched solver has no
    block at all."""
    monkeypatch.setenv("NRGK_MIXED_MAX", "0")
    K, L = 128, 2.5
    V = np.sqrt(2 * 0.02 / np.pi)
    g = nk.NrgCore(nk.Site.spinhalf(-0.25, 0.5), nk.Site.spinhalf(0, 0), K)
    o = oracle_mod.Chain(oracle_mod.Model("spinhalf", eps=-0.25, U=0.5), oracle_mod.Model("spinhalf"), K)
    ts = [V] + [oracle_mod.hopping(n, L) for n in range(5)]
    rs = [1.0] + [1.0 if n == 0 else np.sqrt(L) for n in range(5)]
    for it, (t, r) in enumerate(zip(ts, rs)):
        g.add_bath_site([t, t], r)
        o.add_bath_site([t, t], r)
        g.update_internal_state()
        o.update_internal_state()
        assert np.array_equal(g.kept_counts(), o.kept_counts()), it
        a, b = g.all_eigenvalue, o.all_eigenvalue()
        assert len(a) == len(b) and np.abs(a - b).max() < 1e-8, (it, np.abs(a - b).max())


def test_broadening_matches_oracle(nk, oracle_mod):
    rng = np.random.default_rng(1)
    x = np.sort(10 ** rng.uniform(-9, 1, 300))
    w = rng.uniform(0, 1, 300)
    om = np.logspace(-10, 1, 500)
    a = nk.broaden(x, w, om, 0.6)
    b = oracle_mod.broaden(x, w, om, 0.6)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_plain_fp64_mode_matches_oracle(nk, oracle_mod):
    K, L = 128, 2.5
    V = np.sqrt(2 * 0.02 / np.pi)
    g = nk.NrgCore(nk.Site.spinhalf(-0.25, 0.5), nk.Site.spinhalf(0, 0), K, mode=nk.NRGK_MODE_PLAIN64)
    o = oracle_mod.Chain(oracle_mod.Model("spinhalf", eps=-0.25, U=0.5), oracle_mod.Model("spinhalf"), K)
    ts = [V] + [oracle_mod.hopping(n, L) for n in range(10)]
    rs = [1.0] + [1.0 if n == 0 else np.sqrt(L) for n in range(10)]
    for it, (t, r) in enumerate(zip(ts, rs)):
        g.add_bath_site([t, t], r)
        o.add_bath_site([t, t], r)
        g.update_internal_state()
        o.update_internal_state()
        assert np.array_equal(g.kept_counts(), o.kept_counts()), it
        assert np.abs(g.all_eigenvalue - o.all_eigenvalue()).max() < 1e-8


def test_error_paths(nk):
    g = nk.NrgCore(nk.Site.spinhalf(-0.25, 0.5), nk.Site.spinhalf(0, 0), 64)
    with pytest.raises(nk.NrgkError) as e:
        g.add_bath_site([0.1, 0.1, 0.1, 0.1], 1.0)  # 4 hoppings for 2 operators (nrgcore.hpp:148-152)
    assert "t_hopping" in str(e.value)
    with pytest.raises(nk.NrgkError):
        g.update_internal_state()  # before add_bath_site
    g.add_bath_site([0.1, 0.1], 1.0)
    with pytest.raises(nk.NrgkError):
        g.add_bath_site([0.1, 0.1], 1.0)  # twice without update
    g.update_internal_state()
    with pytest.raises(nk.NrgkError):  # nrgcore::test: operator counts differ (nrgcore.hpp:123-133)
        nk.NrgCore(nk.Site.spinhalf(0, 0).duplicate_ops(), nk.Site.spinhalf(0, 0), 64)


@pytest.mark.gpu
@pytest.mark.parametrize("n_copies", [1, 3])
def test_cpp_nrgcore_api_rgflow_matches_oracle(oracle_mod, n_copies):
    """The C++ driver written against include/nrgk/nrgcore.hpp (examples/rgflowSIAM form) runs the
    chain on the GPU; every iteration's all_eigenvalue equals the oracle's (same state counts, values
    to 1e-9).  With 3 copies the chains are stepped together (nrgk::step_chains) and the driver
    checks that the copies agree bit for bit."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "nrgplusplus_amd", "nrgk_rgflow_siam")
    if not os.path.exists(exe):
        pytest.fail("nrgk_rgflow_siam is not built (run __graft_entry__.build())")
    eps, U, G, L, K, ns = -0.25, 0.5, 0.02, 2.5, 512, 20
    r = subprocess.run([exe, str(eps), str(U), str(G), str(L), str(K), str(ns), "100000", str(n_copies)],
                       capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    assert "diag_blocks ok" in r.stderr  # nrgk::diag_blocks on blocks with known eigenvalues
    lines = [list(map(float, ln.split())) for ln in r.stdout.strip().splitlines()]
    assert len(lines) == ns + 1
    ref = oracle_mod.run_rgflow(oracle_mod.driver(imp_kind=0, eps=eps, U=U, bath_kind=0, Lambda=L,
                                                  first_V=np.sqrt(2 * G / np.pi), nsites=ns, K=K))
    for n in range(ns):
        got = lines[n + 1]
        assert int(got[0]) == n + 1 and int(got[1]) == len(ref[n]), n
        vals = np.array(got[2:-1])
        assert len(vals) == len(ref[n])
        assert np.abs(vals - ref[n]).max() < 1e-9, (n, np.abs(vals - ref[n]).max())


@pytest.mark.gpu
def test_cpp_reference_form_entropy_driver(golden, tmp_path):
    """examples/entropySIAM/main.cpp written against include/nrgk/nrgcore.hpp with the reference's own
    calls in their order (NrgData saveCurrentData / loadCurrentData, a backward loop decrementing
    nrg_iterations_cnt, fdmThermodynamics(&siam, T).calcThermodynamics(scale) / saveFinalData,
    h5stream writes): its HDF5 output against the reference's resultSIAMU0.200000.h5 (entropy to 1e-8,
    specific heat to 2e-6 relative, where the exp clamp does not dominate)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "nrgplusplus_amd", "nrgk_entropy_siam")
    if not os.path.exists(exe):
        pytest.fail("nrgk_entropy_siam is not built (run __graft_entry__.build())")
    out = str(tmp_path / "resultSIAM.h5")
    r = subprocess.run([exe, out], capture_output=True, text=True, timeout=200)
    assert r.returncode == 0, r.stderr
    import nrgplusplus_amd as nk
    gd = golden("siam_U02_L25_K512")
    T = nk.h5_read(out, "temperatureArray")
    S, Cv = nk.h5_read(out, "entropy"), nk.h5_read(out, "specificHeat")
    np.testing.assert_allclose(T, gd["temperatureArray"], rtol=1e-14)
    m = T >= 3e-7
    assert np.abs(S - gd["entropy"])[m].max() < 1e-8
    rel = np.abs(Cv - gd["specificHeat"])[m] / np.maximum(np.abs(gd["specificHeat"][m]), 1e-3)
    assert rel.max() < 2e-6
    assert len(nk.h5_read(out, "Eigenvalues40")) > 0


@pytest.mark.gpu
def test_recorder_load_reproduces_the_backward_pass(nk, tmp_path):
    """NrgData read-back (nrgData.hpp:214-293): the shells a recorder saved are loaded into a recorder
    of a fresh chain of the same models; its FDM backward pass and thermodynamics equal the original's
    (FDM bins summed with atomics: 1e-12 relative; thermodynamics bit-identical)."""
    L, K, ns = 2.0, 128, 16
    def forward():
        g = nk.NrgCore(nk.Site.spinhalf(-0.15, 0.3), nk.Site.spinhalf(0, 0), K)
        op = nk.SystemOperator(g, nk.Site.fdag_chi_spinhalf())
        rec = nk.FdmRecorder(g, op, 400, 1e-10, 10.0)
        return g, op, rec
    g, op, rec = forward()
    V = np.sqrt(2 * 0.01 / np.pi)
    g.add_bath_site([V, V], 1.0)
    op.update()
    g.update_internal_state()
    m = 0
    for n in range(ns):
        t = nk.hopping(n, L)
        g.add_bath_site([t, t], 1.0 if n == 0 else np.sqrt(L))
        op.update()
        if g.checkHigherEnergyDiscarded():
            rec.record()
        else:
            m += 1
        g.update_internal_state()
    path = str(tmp_path / "shells.h5")
    rec.save(path)
    T = np.logspace(-6, 0, 40)
    a = rec.backward(L, scale_offset=-1, first_cnt=m + 1)
    ta = rec.thermodynamics(L, T, scale_offset=-1, first_cnt=m + 1)
    g2, op2, rec2 = forward()
    rec2.load(path)
    b = rec2.backward(L, scale_offset=-1, first_cnt=m + 1)
    tb = rec2.thermodynamics(L, T, scale_offset=-1, first_cnt=m + 1)
    for x, y in zip(a[:2], b[:2]):
        assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max()
    np.testing.assert_allclose(a[2], b[2], rtol=1e-13)
    for x, y in zip(ta, tb):
        np.testing.assert_array_equal(x, y)
