"""libnrgk.so without a GPU: it loads, exports every entry point include/nrgk.h declares, and its
host-side model construction matches the oracle (which restates fermionBasis / spinhalf /
twoChannel / spinhalfKondo).  No device calls."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "nrgk.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*|void)\s+(nrgk_[a-z0-9_]+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol(nk):
    lib = os.path.join(ROOT, "nrgplusplus_amd", "libnrgk.so")
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (nrgk_[a-z0-9_]+)$", out, re.M))
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing
    assert len(_declared()) >= 40


def test_library_is_built_for_gfx950():
    lib = os.path.join(ROOT, "nrgplusplus_amd", "libnrgk.so")
    out = subprocess.run(["strings", lib], capture_output=True, text=True, check=True).stdout
    assert "gfx950" in out


@pytest.mark.parametrize("which", ["spinhalf", "two_channel", "kondo", "spinhalf_h"])
def test_host_models_match_oracle(nk, oracle_mod, which):
    if which == "spinhalf":
        a, b = nk.Site.spinhalf(-0.15, 0.3), oracle_mod.Model("spinhalf", eps=-0.15, U=0.3)
    elif which == "spinhalf_h":
        a, b = nk.Site.spinhalf(-0.1, 0.7, 0.013), oracle_mod.Model("spinhalf", eps=-0.1, U=0.7, h=0.013)
    elif which == "two_channel":
        a, b = nk.Site.two_channel(), oracle_mod.Model("two_channel")
    else:
        a, b = nk.Site.kondo(0.25, 0.5), oracle_mod.Model("kondo", J=0.25, S=0.5)
    qa, da, ea, ca, oa = a.export()
    qb, db, eb, cb, ob = b.export()
    assert np.array_equal(qa, qb) and np.array_equal(da, db) and np.array_equal(ca, cb)
    if which == "kondo":  # 2x2 blocks diagonalised by different small solvers
        assert np.abs(ea - eb).max() < 1e-14
    else:
        assert np.array_equal(ea, eb)  # bit-exact
    assert len(oa) == len(ob)
    for pa, pb in zip(oa, ob):
        assert [(i, j) for i, j, _ in pa] == [(i, j) for i, j, _ in pb]
        for (_, _, m1), (_, _, m2) in zip(pa, pb):
            if which == "kondo":
                # eigenvector signs are a gauge: compare |elements|
                assert np.abs(np.abs(m1) - np.abs(m2)).max() < 1e-14
            else:
                assert np.array_equal(m1, m2)


def _full(site):
    """Full Fock-space matrices of the f-dagger operators from the block form."""
    qn, dims, _, _, ops = site.export()
    off = np.concatenate([[0], np.cumsum(dims)])
    D = off[-1]
    mats = []
    for blocks in ops:
        M = np.zeros((D, D))
        for i, j, b in blocks:
            M[off[i]:off[i + 1], off[j]:off[j + 1]] = b
        mats.append(M)
    return mats


@pytest.mark.parametrize("which", ["spinhalf", "two_channel"])
def test_jordan_wigner_anticommutators(nk, which):
    site = nk.Site.spinhalf(0.0, 0.0) if which == "spinhalf" else nk.Site.two_channel()
    f = _full(site)
    D = f[0].shape[0]
    for i in range(len(f)):
        for j in range(len(f)):
            acomm = f[i].T @ f[j] + f[j] @ f[i].T  # {f_i, f_j^dagger}
            np.testing.assert_array_equal(acomm, np.eye(D) if i == j else np.zeros((D, D)))
            np.testing.assert_array_equal(f[i] @ f[j] + f[j] @ f[i], np.zeros((D, D)))


def test_fdag_chi_observable(nk, oracle_mod):
    """f^dagger chi blocks: chi = (-1)^N flips the sign of f^dagger acting on odd states."""
    s = nk.Site.fdag_chi_spinhalf()
    f = _full(s)
    base = _full(nk.Site.spinhalf(0.0, 0.0))
    qn, dims, _, _, _ = s.export()
    n_el = np.repeat(qn.sum(axis=1), dims)
    chi = np.diag((-1.0) ** n_el)
    for a, b in zip(f, base):
        np.testing.assert_array_equal(a, b @ chi)


def test_double_occupancy_observable(nk):
    """n_up n_dn (spinhalf.hpp:106-114) from the f-dagger matrices of the same basis: f f^T products."""
    f = _full(nk.Site.spinhalf(0.0, 0.0))
    d = _full(nk.Site.double_occupancy_spinhalf())
    assert len(d) == 1
    np.testing.assert_array_equal(d[0], (f[0] @ f[0].T) @ (f[1] @ f[1].T))
    qn, dims, _, _, ops = nk.Site.double_occupancy_spinhalf().export()
    # one 1x1 block, in the doubly occupied sector
    assert [(i, j) for i, j, _ in ops[0]] == [(int(np.nonzero((qn == [1, 1]).all(1))[0][0]),) * 2]


def test_duplicate_ops(nk):
    s = nk.Site.spinhalf(-0.2, 0.4).duplicate_ops()
    _, _, _, _, ops = s.export()
    assert len(ops) == 4
    for p in (0, 1):
        for (i, j, m), (i2, j2, m2) in zip(ops[p], ops[p + 2]):
            assert (i, j) == (i2, j2) and np.array_equal(m, m2)


def test_cpp_nrgcore_mirror_compiles():
    """The C++ host API (include/nrgk/nrgcore.hpp, the reference's nrgcore method names) and its
    RG-flow driver compile as plain host C++ against the C ABI."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = os.path.join(root, "nrgplusplus_amd", "cpp", "rgflow_siam.cpp")
    r = subprocess.run(["g++", "-fsyntax-only", "-std=c++17", "-Wall", "-Werror", src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


# ---------------------------------------------------------------- other model families
def _jw(n_orb):
    """f^dagger_i with orbital i as bit i of the Fock index and the sign of the occupied orbitals
    below i (the reference's Jordan-Wigner order, fermionBasis.hpp:397-420)"""
    D = 1 << n_orb
    ops = []
    for i in range(n_orb):
        m = np.zeros((D, D))
        for s in range(D):
            if not (s >> i) & 1:
                m[s | (1 << i), s] = (-1) ** bin(s & ((1 << i) - 1)).count("1")
        ops.append(m)
    return ops


@pytest.mark.parametrize("sym", ["charge_and_spin", "charge_only", "spin_only"])
def test_fock_basis(nk, sym):
    f, q = nk.fock_basis(3, sym)
    for a, b in zip(f, _jw(3)):
        np.testing.assert_array_equal(a, b)
    occ = np.array([[(s >> i) & 1 for i in range(3)] for s in range(8)])
    expect = {"charge_and_spin": np.stack([occ[:, 0] + occ[:, 2], occ[:, 1]], 1),
              "charge_only": occ.sum(1)[:, None],
              "spin_only": (occ[:, 0] + occ[:, 2] - occ[:, 1])[:, None]}[sym]
    np.testing.assert_array_equal(q, expect)


def _check_site(site, qn_state, H, ops):
    """sectors in first-appearance order, per-sector eigenvalues of H (numpy), chi = (-1)^sum(qn), and
    every operator block's Frobenius norm equal to its unrotated block's (orthogonal rotation)"""
    qn, dims, en, chi, got_ops = site.export()
    keys = []
    for q in map(tuple, qn_state):
        if q not in keys:
            keys.append(q)
    np.testing.assert_array_equal(qn, np.array(keys).reshape(len(keys), -1))
    idx = [[s for s in range(len(qn_state)) if tuple(qn_state[s]) == k] for k in keys]
    np.testing.assert_array_equal(dims, [len(x) for x in idx])
    np.testing.assert_array_equal(chi, [(-1.0) ** sum(k) for k in keys])
    off = 0
    for x in idx:
        np.testing.assert_allclose(en[off:off + len(x)], np.linalg.eigvalsh(H[np.ix_(x, x)]), atol=1e-14)
        off += len(x)
    assert len(got_ops) == len(ops)
    for blocks, F in zip(got_ops, ops):
        want = {(i, j): np.linalg.norm(F[np.ix_(idx[i], idx[j])]) for i in range(len(idx)) for j in range(len(idx))}
        assert sorted((i, j) for i, j, _ in blocks) == sorted(k for k, v in want.items() if v > 0)
        for i, j, b in blocks:
            assert abs(np.linalg.norm(b) - want[(i, j)]) < 1e-13


def test_reference_models(nk):
    f3, _ = nk.fock_basis(3, "charge_only")
    f1, f2 = _jw(1), _jw(2)
    q1 = np.array([[(s & 1)] for s in range(2)])
    _check_site(nk.Site.spinless(0.3), q1, 0.3 * f1[0] @ f1[0].T, f1)
    q2 = np.array([[bin(s).count("1")] for s in range(4)])
    _check_site(nk.Site.spinless_two_lead(-0.2), q2, -0.2 * f2[0] @ f2[0].T, f2)
    # resonantTwoLead.hpp:31-40: eps n0 + V_L(f0+ f1 + h.c.) + V_R(f0+ f2 + h.c.) + U (n0 - 1)(n1 + n2 - 1)
    eps, vl, vr, U = -0.025, 0.11, 0.07, 0.05
    f = _jw(3)
    n = [x @ x.T for x in f]
    I = np.eye(8)
    H = eps * n[0] + vl * (f[0] @ f[1].T + f[1] @ f[0].T) + vr * (f[0] @ f[2].T + f[2] @ f[0].T) \
        + U * (n[0] - I) @ (n[1] + n[2] - I)
    q3 = np.array([[bin(s).count("1")] for s in range(8)])
    _check_site(nk.Site.resonant_two_lead(eps, vl, vr, U), q3, H, [f[1], f[2]])
    # impurityFdag and impurityNparticle = sum_j F_ij F_ij^T on (i, i) (resonantTwoLead.hpp:63-74)
    _check_site(nk.Site.resonant_two_lead(eps, vl, vr, U, observables=True), q3, H, [f[0], n[0]])
    # spinhalfSC with pairing: sector Sz = 0 holds |0> and |up dn>
    e, u, h, d = -0.1, 0.3, 0.02, 0.05
    nu, nd = f2[0] @ f2[0].T, f2[1] @ f2[1].T
    Hs = e * (nu + nd) + h * (nu - nd) + u * nu @ nd + d * (f2[0].T @ f2[1].T + f2[1] @ f2[0])
    qs = np.array([[(s & 1) - ((s >> 1) & 1)] for s in range(4)])
    _check_site(nk.Site.spinhalf_sc(e, u, h, d), qs, Hs, f2)
    _check_site(nk.Site.free("spin_only"), qs, np.zeros((4, 4)), f2)
    # a user model through the generic builder equals the library's own
    _check_site(nk.Site.from_hamiltonian(q3, H, [f[1], f[2]]), q3, H, [f[1], f[2]])


def test_model_builder_rejects_symmetry_breaking(nk):
    f = _jw(2)
    q = np.array([[bin(s).count("1")] for s in range(4)])
    with pytest.raises(nk.NrgkError):  # f0+ + f0 changes N
        nk.Site.from_hamiltonian(q, f[0] + f[0].T, [f[0]])
    with pytest.raises(nk.NrgkError):  # not symmetric
        nk.Site.from_hamiltonian(q, f[0] @ f[1].T, [f[0]])


def test_hdf5_reader_errors(nk, tmp_path):
    """the HDF5 library loads on demand; a missing file or dataset is an error, not a crash"""
    with pytest.raises(nk.NrgkError):
        nk.h5_read(str(tmp_path / "absent.h5"), "/x")


def test_chains_step_argument_checks(nk):
    """nrgk_chains_step rejects bad arguments before any chain is touched (no device needed): null
    chain list or chain, negative sizes; an empty list is a no-op."""
    import ctypes as C
    lib = nk.lib()
    one = np.ones(2)
    r = np.ones(1)
    assert lib.nrgk_chains_step(None, 0, None, 0, None, None, 0, None, 0, None) == 0
    assert lib.nrgk_chains_step(None, 1, nk._dp(one), 2, nk._dp(r), None, 0, None, 0, None) != 0
    nulls = (C.c_void_p * 1)(None)
    assert lib.nrgk_chains_step(nulls, 1, nk._dp(one), 2, nk._dp(r), None, 0, None, 0, None) != 0
    assert b"null chain" in lib.nrgk_last_error()
    assert lib.nrgk_chains_step(nulls, -1, nk._dp(one), 2, nk._dp(r), None, 0, None, 0, None) != 0


def test_eigh_batched_argument_checks(nk):
    """nrgk_eigh_batched rejects a null context, negative sizes and an unknown mode before touching a
    device"""
    lib = nk.lib()
    dims = np.array([2], np.int32)
    H = np.eye(2)
    w, U = np.zeros(2), np.zeros(4)
    assert lib.nrgk_eigh_batched(None, 1, nk._ip(dims), nk._dp(H), nk._dp(w), nk._dp(U), nk.NRGK_MODE_MIXED) != 0
    assert lib.nrgk_eigh_batched(None, -1, None, None, None, None, nk.NRGK_MODE_MIXED) != 0
