// z-shifted logarithmic discretisation of a bath and its Wilson chain (the reference's z-averaging
// generator, independFiles/zAverage/hoppingElement.cpp:9-105 for the flat band and
// independFiles/zAverage/zAverage.hpp:12-195 for a given density of states).
//
// The band [-D, D] is cut at D Lambda^{-(i + z - 1)} (i >= 1) into the intervals
// I_0 = [D Lambda^{-z}, D] and I_i = [D Lambda^{-(i + z)}, D Lambda^{-(i + z - 1)}], mirrored at
// negative energies (zAverage.hpp:168-176).  Each interval becomes one level at the Campo-Oliveira
// energy xi = int_I rho / int_I rho/e with coupling sqrt(int_I rho) to the first Wilson site
// (zAverage.hpp:90-130; for the flat band these are hoppingElement.cpp:9-15 and :33-38).  The
// tridiagonal (Wilson) chain of that star follows by Lanczos (zAverage.hpp:59-77).
//
// The reference runs a plain three-term Lanczos in 1000-digit MPFR arithmetic.  Here it runs in
// quadruple precision (113-bit significand) with full reorthogonalisation (classical Gram-Schmidt,
// twice), which keeps the Krylov basis orthogonal to working precision; the hoppings t_n fall like
// Lambda^{-n/2}, and a basis vector of site n lives at energies of that order, so their relative
// accuracy stays near 1e-30 times the ratio of the band edge to that scale.
//
// Host code only (no GPU work): compiled by the host compiler, linked into libnrgk.
#include <quadmath.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nrgk.h"

namespace {

using q128 = __float128;

// integrals over [a, b] (0 < a < b or a < b < 0) of a tabulated density, linear between its points
// (and zero outside them): int rho and int rho / e
struct Tab {
  const double* x;
  const double* r;
  int n;
  double at(double e) const {
    if (e < x[0] || e > x[n - 1]) return 0.0;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
      const int m = (lo + hi) / 2;
      (x[m] <= e ? lo : hi) = m;
    }
    const double w = x[hi] > x[lo] ? (e - x[lo]) / (x[hi] - x[lo]) : 0.0;
    return r[lo] + w * (r[hi] - r[lo]);
  }
  // exact integrals of the piecewise-linear interpolant
  void exact(q128 a, q128 b, q128& I0, q128& I1) const {
    I0 = 0;
    I1 = 0;
    for (int k = 0; k + 1 < n; k++) {
      const q128 x0 = x[k], x1 = x[k + 1];
      const q128 lo = a > x0 ? a : x0, hi = b < x1 ? b : x1;
      if (!(hi > lo) || !(x1 > x0)) continue;
      const q128 beta = ((q128)r[k + 1] - (q128)r[k]) / (x1 - x0), alpha = (q128)r[k] - beta * x0;
      I0 += alpha * (hi - lo) + beta * (hi * hi - lo * lo) / 2;
      I1 += alpha * logq(hi / lo) + beta * (hi - lo);  // lo, hi of one sign
    }
  }
  // the reference's rule (zAverage/quadrature.hpp:13-26): 100 left-point samples
  void left_sum(q128 a, q128 b, q128& I0, q128& I1) const {
    const int m = 100;
    const q128 dx = (b - a) / m;
    I0 = 0;
    I1 = 0;
    for (int i = 0; i < m; i++) {
      const q128 e = a + i * dx;
      const q128 v = at((double)e);
      I0 += v;
      I1 += v / e;
    }
    I0 *= dx;
    I1 *= dx;
  }
};

}  // namespace

extern "C" int nrgk_zdisc_chain(double z, double lambda, double D, const double* x, const double* rho, int n_x,
                                int quad, int n_intervals, int normalize, int n_sites, double* t_out,
                                double* eps_out) {
  if (!(lambda > 1.0) || !(D > 0.0) || !(z >= 0.0) || n_sites <= 0 || !t_out || !eps_out || quad < 0 || quad > 1)
    return NRGK_ERR_INVALID;
  if (n_x != 0 && (n_x < 2 || !x || !rho)) return NRGK_ERR_INVALID;
  for (int k = 0; k + 1 < n_x; k++)
    if (!(x[k + 1] >= x[k])) return NRGK_ERR_INVALID;
  // z and z + 1 give the same mesh (I_0 of z = 0 is empty): z = 0 is taken as z = 1
  const q128 Z = z == 0.0 ? (q128)1 : (q128)z;
  const q128 L = lambda, lnL = logq(L), Dq = D;
  const int M = n_intervals > 0 ? n_intervals : 2 * n_sites + 20;
  // the star: levels xi_k and couplings g_k, negative side then positive side
  std::vector<q128> xi(2 * M), g(2 * M);
  const Tab tab{x, rho, n_x};
  for (int i = 0; i < M; i++) {
    q128 e, w;  // positive-side level and weight
    if (n_x == 0) {  // flat band rho = 1 / (2D): hoppingElement.cpp:9-15, 33-38
      if (i == 0) {
        e = Dq * (1 - powq(L, -Z)) / (Z * lnL);
        w = (1 - powq(L, -Z)) / 2;
      } else {
        e = Dq * (1 - 1 / L) * powq(L, 1 - i - Z) / lnL;
        w = (1 - 1 / L) * powq(L, 1 - i - Z) / 2;
      }
      xi[M + i] = e;
      g[M + i] = sqrtq(w);
      xi[M - 1 - i] = -e;
      g[M - 1 - i] = sqrtq(w);
    } else {
      const q128 hi = i == 0 ? Dq : Dq * powq(L, 1 - i - Z), lo = Dq * powq(L, -(i + Z));
      q128 I0, I1;
      for (int side = 0; side < 2; side++) {
        const q128 a = side ? lo : -hi, b = side ? hi : -lo;
        if (quad == 0) tab.exact(a, b, I0, I1);
        else tab.left_sum(a, b, I0, I1);
        const int k = side ? M + i : M - 1 - i;
        xi[k] = I1 != 0 ? I0 / I1 : (a + b) / 2;
        g[k] = sqrtq(fabsq(I0));
      }
    }
  }
  const int m = 2 * M;
  // Lanczos from the coupling vector (normalised unless asked otherwise: hoppingElement.cpp:45-53
  // leaves it as is, zAverage.hpp:131-143 normalises it)
  std::vector<std::vector<q128>> V;
  std::vector<q128> v = g, r(m), prev(m, 0);
  q128 nrm = 0;
  for (auto a : v) nrm += a * a;
  if (!(nrm > 0)) return NRGK_ERR_INVALID;
  if (normalize)
    for (auto& a : v) a /= sqrtq(nrm);
  q128 t_prev = 0;
  const int steps = n_sites < m ? n_sites : m - 1;
  for (int n = 0; n < n_sites; n++) {
    if (n >= steps) {  // the star has m levels: the chain ends after m - 1 hoppings
      t_out[n] = 0.0;
      eps_out[n] = 0.0;
      continue;
    }
    q128 a = 0;
    for (int k = 0; k < m; k++) a += xi[k] * v[k] * v[k];
    for (int k = 0; k < m; k++) r[k] = xi[k] * v[k] - a * v[k] - t_prev * prev[k];
    V.push_back(v);
    for (int pass = 0; pass < 2; pass++)
      for (auto& u : V) {
        q128 c = 0, uu = 0;
        for (int k = 0; k < m; k++) {
          c += r[k] * u[k];
          uu += u[k] * u[k];
        }
        c /= uu;  // the first vector may be unnormalised
        for (int k = 0; k < m; k++) r[k] -= c * u[k];
      }
    q128 t = 0;
    for (int k = 0; k < m; k++) t += r[k] * r[k];
    t = sqrtq(t);
    const q128 sc = powq(L, (q128)n / 2);
    t_out[n] = (double)(t * sc);
    eps_out[n] = (double)(a * sc);
    prev = v;
    for (int k = 0; k < m; k++) v[k] = t > 0 ? r[k] / t : 0;
    t_prev = t;
  }
  return NRGK_OK;
}
