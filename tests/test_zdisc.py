"""z-shifted logarithmic discretisation and its Wilson chain (host code of libnrgk, no GPU).

Reference: independFiles/zAverage/hoppingElement.cpp:9-105 (calcTn, flat band) and
independFiles/zAverage/zAverage.hpp:12-195 (tabulated density of states).

Pins:
  * z = 1, flat band: the chain is Wilson's closed form divided by A_Lambda = ln(L)(1+1/L)/(2(1-1/L)),
    t_n Lambda^{n/2} A_Lambda = hopping(n, Lambda) (examples/rgflowSIAM/main.cpp:5-10), eps_n = 0;
  * the reference's own output hoppingElementsL3.h5 (calcTn, z = 1, Lambda = 3, 11 intervals): its
    first entries, before its three-term recurrence with an unnormalised start vector breaks down;
  * any density: the chain's site-0 moments (T^j)_00 equal the moments of the discretised star
    (Gauss quadrature property of Lanczos), with the star's levels and weights computed here by
    Gauss-Legendre integration of the tabulated density;
  * a flat density given as a table equals the flat closed form; z = 0 is the z = 1 mesh.
"""
import math

import numpy as np
import pytest


def _A(L):
    return 0.5 * math.log(L) * (1 + 1 / L) / (1 - 1 / L)


@pytest.mark.parametrize("L", [2.0, 2.5, 3.0, 4.0])
def test_flat_band_z1_is_wilson_chain(nk, L):
    t, e = nk.z_chain(1.0, L, 60)
    w = np.array([nk.hopping(n, L) for n in range(60)])
    np.testing.assert_allclose(t * _A(L), w, rtol=1e-13)
    assert np.abs(e).max() < 1e-25


def test_reference_calcTn_output(nk, golden):
    """hoppingElementsL3.h5 (`tnZ1.000lmd3.000`): calcTn with hcmax = 10 (11 intervals per side) and
    the coupling vector left unnormalised.  Entries 0-3 agree to 2e-10 and entry 4 to 1e-8; from
    there on the reference's plain recurrence, started from a vector of norm below 1 (the mesh
    stops at Lambda^{-11}), drifts (1.5e-6 at 5, 1e-3 at 6) and breaks down at entry 7 (1.447,
    44.6, ...), while the reorthogonalised chain of the same 22-level star stays regular."""
    ref = golden("zdisc_flat_z1_L3")["tnZ1.000lmd3.000"]
    t, _ = nk.z_chain(1.0, 3.0, 12, n_intervals=11, normalize=False)
    np.testing.assert_allclose(t[:4], ref[:4], rtol=2e-10)
    np.testing.assert_allclose(t[4], ref[4], rtol=1e-8)
    assert ref[7] > 1.0 and t[7] < 0.61  # the reference's breakdown, not reproduced


def test_z0_is_z1_and_tabulated_flat_equals_closed_form(nk):
    t1, _ = nk.z_chain(1.0, 2.5, 40)
    t0, _ = nk.z_chain(0.0, 2.5, 40)
    np.testing.assert_array_equal(t0, t1)
    for z in (0.37, 1.0):
        tf, ef = nk.z_chain(z, 2.5, 40, D=2.0)
        tt, et = nk.z_chain(z, 2.5, 40, D=2.0, dos=([-2.0, 2.0], [0.25, 0.25]))
        np.testing.assert_allclose(tt, tf, rtol=1e-13)
        assert np.abs(et).max() < 1e-12 * np.abs(tt).max()


def _star(z, L, D, x, r, M):
    """levels xi = int rho / int rho/e and weights int rho of the z-shifted intervals, by 24-point
    Gauss-Legendre on every piece of the piecewise-linear density"""
    gx, gw = np.polynomial.legendre.leggauss(24)

    def integrals(a, b):
        cuts = np.unique(np.clip(np.concatenate([[a, b], x]), a, b))
        i0 = i1 = 0.0
        for lo, hi in zip(cuts[:-1], cuts[1:]):
            if hi <= lo:
                continue
            e = 0.5 * (hi - lo) * gx + 0.5 * (hi + lo)
            v = np.interp(e, x, r, left=0.0, right=0.0)
            i0 += 0.5 * (hi - lo) * np.sum(gw * v)
            i1 += 0.5 * (hi - lo) * np.sum(gw * v / e)
        return i0, i1

    xi, w = [], []
    for i in range(M):
        hi = D if i == 0 else D * L ** (1 - i - z)
        lo = D * L ** (-(i + z))
        for a, b in ((lo, hi), (-hi, -lo)):
            i0, i1 = integrals(a, b)
            xi.append(i0 / i1)
            w.append(i0)
    return np.array(xi), np.array(w)


@pytest.mark.parametrize("kind", ["semicircle", "asymmetric"])
def test_chain_reproduces_moments_of_the_discretised_density(nk, kind):
    z, L, D, n = 0.6, 3.0, 1.0, 8
    x = np.linspace(-D, D, 401)
    r = np.sqrt(np.maximum(0.0, 1 - x ** 2)) if kind == "semicircle" else 0.5 * (1 + 0.6 * x) * (1 - x ** 2)
    M = 2 * n + 20
    t, e = nk.z_chain(z, L, n, D=D, dos=(x, r), n_intervals=M)
    sc = L ** (np.arange(n) / 2)
    T = np.diag(e / sc) + np.diag((t / sc)[: n - 1], 1) + np.diag((t / sc)[: n - 1], -1)
    xi, w = _star(z, L, D, x, r, M)
    w = w / w.sum()
    for j in range(2 * n):  # (T^j)_00 = sum_k w_k xi_k^j for j <= 2n - 1
        mu = np.sum(w * xi ** j)
        got = np.linalg.matrix_power(T, j)[0, 0]
        assert abs(got - mu) <= 1e-11 * max(1.0, abs(mu)), (j, got, mu)
    if kind == "semicircle":
        assert np.abs(e).max() < 1e-12  # symmetric density: no on-site energies


def test_reference_quadrature_rule(nk):
    """quad="reference" samples the density at 100 left points per interval (zAverage/quadrature.hpp:
    13-26); for a linear density the first level and weight follow from those sums by hand."""
    z, L = 0.5, 2.0
    x, r = np.array([-1.0, 1.0]), np.array([0.2, 0.8])
    t, e = nk.z_chain(z, L, 1, dos=(x, r), quad="reference", n_intervals=1)
    # one interval per side: the star has two levels and the chain one hopping
    lv, wt = [], []
    for a, b in ((L ** -z, 1.0), (-1.0, -(L ** -z))):
        dx = (b - a) / 100
        s = a + dx * np.arange(100)
        rho = np.interp(s, x, r)
        i0, i1 = dx * rho.sum(), dx * (rho / s).sum()
        lv.append(i0 / i1)
        wt.append(abs(i0))
    lv, wt = np.array(lv), np.array(wt) / sum(wt)
    a0 = np.sum(wt * lv)
    np.testing.assert_allclose(e[0], a0, rtol=1e-13)
    np.testing.assert_allclose(t[0], math.sqrt(np.sum(wt * (lv - a0) ** 2)), rtol=1e-13)


def test_invalid_arguments(nk):
    with pytest.raises(nk.NrgkError):
        nk.z_chain(0.5, 1.0, 10)
    with pytest.raises(nk.NrgkError):
        nk.z_chain(-0.1, 2.0, 10)
    with pytest.raises(nk.NrgkError):
        nk.z_chain(0.5, 2.0, 10, dos=([1.0, -1.0], [0.5, 0.5]))
