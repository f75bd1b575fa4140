// Site models (host-side setup, not the hot path).
//
// Restates the reference's model construction:
//   fermionBasis (nrgcore/include/models/fermionBasis.hpp:20-442): Jordan-Wigner
//     f_i^dagger = I^(dof-i-1) (x) f^dagger (x) sigma_z^i, so orbital i is bit i of the Fock index
//     and <s|f_i^dagger|s'> = (-1)^(number of occupied orbitals below i);
//     chargeAndSpin quantum numbers [N(even orbitals), N(odd orbitals)] (:122-141);
//     sectors in first-appearance order (:230-268); a block is kept iff sum |elem| > 0 (:291-318).
//   spinhalf (nrgcore/include/models/spinhalf.hpp:81-124), twoChannel
//   (nrgcore/src/models/twoChannel.cpp:12-48), spinhalfKondo (nrgcore/include/models/spinhalfKondo.hpp:28-166)
//   and the f^dagger chi observable of examples/fdmSpectrumSiam/main.cpp:86-106.
// Small blocks are diagonalised with a cyclic FP64 Jacobi (eigenvalues ascending, vectors in columns).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <vector>

#include "common.hpp"

namespace nrgk {
namespace {

struct M {
  int r = 0, c = 0;
  std::vector<double> a;
  M() = default;
  M(int rr, int cc) : r(rr), c(cc), a((size_t)rr * cc, 0.0) {}
  double& operator()(int i, int j) { return a[(size_t)i * c + j]; }
  double operator()(int i, int j) const { return a[(size_t)i * c + j]; }
};
M eye(int n) {
  M m(n, n);
  for (int i = 0; i < n; i++) m(i, i) = 1.0;
  return m;
}
M tr(const M& x) {
  M t(x.c, x.r);
  for (int i = 0; i < x.r; i++)
    for (int j = 0; j < x.c; j++) t(j, i) = x(i, j);
  return t;
}
M mul(const M& x, const M& y) {
  M z(x.r, y.c);
  for (int i = 0; i < x.r; i++)
    for (int k = 0; k < x.c; k++) {
      const double v = x(i, k);
      if (v == 0.0) continue;
      for (int j = 0; j < y.c; j++) z(i, j) += v * y(k, j);
    }
  return z;
}
M kron(const M& x, const M& y) {
  M z(x.r * y.r, x.c * y.c);
  for (int r = 0; r < x.r; r++)
    for (int s = 0; s < x.c; s++)
      for (int v = 0; v < y.r; v++)
        for (int w = 0; w < y.c; w++) z(r * y.r + v, s * y.c + w) = x(r, s) * y(v, w);
  return z;
}
M lin(const M& x, double a, const M& y, double b) {
  M z(x.r, x.c);
  for (size_t i = 0; i < z.a.size(); i++) z.a[i] = x.a[i] * a + y.a[i] * b;
  return z;
}
M scal(const M& x, double a) {
  M z(x.r, x.c);
  for (size_t i = 0; i < z.a.size(); i++) z.a[i] = x.a[i] * a;
  return z;
}

// cyclic Jacobi for a small symmetric block: returns ascending eigenvalues, U with vectors in columns
std::vector<double> small_eigh(const M& A, M& U) {
  const int n = A.r;
  M a = A;
  U = eye(n);
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += a(p, q) * a(p, q);
    if (off == 0.0) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        if (a(p, q) == 0.0) continue;
        const double th = (a(q, q) - a(p, p)) / (2 * a(p, q));
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double ukp = U(k, p), ukq = U(k, q);
          U(k, p) = c * ukp - s * ukq;
          U(k, q) = s * ukp + c * ukq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; i++) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  std::vector<double> w(n);
  M Us(n, n);
  for (int j = 0; j < n; j++) {
    w[j] = a(ord[j], ord[j]);
    for (int i = 0; i < n; i++) Us(i, j) = U(i, ord[j]);
  }
  U = Us;
  return w;
}

struct Basis {
  std::vector<M> fdag;                   // full f_i^dagger
  std::vector<std::vector<double>> occ;  // occupation of orbital i per Fock state
  std::vector<std::vector<int>> qn_state, unique_qn;
  std::vector<std::vector<int>> blocks;
  void jordan_wigner(int dof) {
    const int D = 1 << dof;
    fdag.assign(dof, M(D, D));
    occ.assign(dof, std::vector<double>(D, 0.0));
    for (int i = 0; i < dof; i++)
      for (int s = 0; s < D; s++) {
        occ[i][s] = (s >> i) & 1;
        if ((s >> i) & 1) continue;
        const int sign = (__builtin_popcount(s & ((1 << i) - 1)) & 1) ? -1 : 1;
        fdag[i](s | (1 << i), s) = sign;
      }
  }
  void charge_and_spin() {
    qn_state.clear();
    for (size_t s = 0; s < occ[0].size(); s++) {
      int up = 0, dn = 0;
      for (size_t j = 0; j < occ.size(); j++) (j % 2 == 0 ? up : dn) += (int)occ[j][s];
      qn_state.push_back({up, dn});
    }
    std::set<std::vector<int>> seen;
    unique_qn.clear();
    blocks.clear();
    for (size_t s = 0; s < qn_state.size(); s++)
      if (seen.insert(qn_state[s]).second) {
        unique_qn.push_back(qn_state[s]);
        std::vector<int> b{(int)s};
        for (size_t t = s + 1; t < qn_state.size(); t++)
          if (qn_state[t] == qn_state[s]) b.push_back((int)t);
        blocks.push_back(b);
      }
  }
  std::vector<std::map<std::pair<int, int>, M>> block_ops(const std::vector<M>& ops) const {
    std::vector<std::map<std::pair<int, int>, M>> out(ops.size());
    for (size_t i = 0; i < blocks.size(); i++)
      for (size_t j = 0; j < blocks.size(); j++)
        for (size_t p = 0; p < ops.size(); p++) {
          M f((int)blocks[i].size(), (int)blocks[j].size());
          double tv = 0;
          for (size_t a = 0; a < blocks[i].size(); a++)
            for (size_t b = 0; b < blocks[j].size(); b++) {
              f((int)a, (int)b) = ops[p](blocks[i][a], blocks[j][b]);
              tv += std::fabs(f((int)a, (int)b));
            }
          if (tv > 0) out[p][{(int)i, (int)j}] = f;
        }
    return out;
  }
  std::vector<M> block_diag(const M& H) const {
    std::vector<M> hb;
    for (auto& b : blocks) {
      M f((int)b.size(), (int)b.size());
      for (size_t a = 0; a < b.size(); a++)
        for (size_t c = 0; c < b.size(); c++) f((int)a, (int)c) = H(b[a], b[c]);
      hb.push_back(f);
    }
    return hb;
  }
};

}  // namespace

// A site owned by the library, viewable as nrgk_site_desc.
struct Site {
  std::vector<int> qn, dims, nblocks, bi, bj;
  std::vector<double> energies, chi, data;
  int n_sectors = 0, n_qn = 2, n_ops = 0;
  nrgk_site_desc view() const {
    nrgk_site_desc d{};
    d.n_sectors = n_sectors;
    d.n_qn = n_qn;
    d.qn = qn.data();
    d.dims = dims.data();
    d.energies = energies.data();
    d.chi = chi.data();
    d.n_ops = n_ops;
    d.op_nblocks = nblocks.data();
    d.op_bi = bi.data();
    d.op_bj = bj.data();
    d.op_data = data.data();
    return d;
  }
};

static Site make_site(const Basis& B, const std::vector<std::vector<double>>& E,
                      const std::vector<std::map<std::pair<int, int>, M>>& ops) {
  Site s;
  s.n_sectors = (int)B.unique_qn.size();
  for (auto& q : B.unique_qn) {
    s.qn.insert(s.qn.end(), q.begin(), q.end());
    int tot = 0;
    for (int x : q) tot += x;
    s.chi.push_back(std::pow(-1., (double)tot));
  }
  for (auto& e : E) {
    s.dims.push_back((int)e.size());
    s.energies.insert(s.energies.end(), e.begin(), e.end());
  }
  s.n_ops = (int)ops.size();
  for (auto& op : ops) {
    s.nblocks.push_back((int)op.size());
    for (auto& [k, m] : op) {
      s.bi.push_back(k.first);
      s.bj.push_back(k.second);
      s.data.insert(s.data.end(), m.a.begin(), m.a.end());
    }
  }
  return s;
}

Site* site_spinhalf(double eps, double U, double h) {
  Basis B;
  B.jordan_wigner(2);
  B.charge_and_spin();
  const M nu = mul(B.fdag[0], tr(B.fdag[0])), nd = mul(B.fdag[1], tr(B.fdag[1]));
  // H = (n_up + n_dn) eps + (n_up - n_dn) h + (n_up n_dn) U, evaluated as in spinhalf.hpp:109-111
  const M H = lin(lin(lin(nu, 1.0, nd, 1.0), eps, lin(nu, 1.0, nd, -1.0), h), 1.0, mul(nu, nd), U);
  std::vector<std::vector<double>> E;
  for (auto& hb : B.block_diag(H)) {
    M Ub;
    E.push_back(small_eigh(hb, Ub));
  }
  return new Site(make_site(B, E, B.block_ops(B.fdag)));
}

Site* site_two_channel() {
  Basis B;
  B.jordan_wigner(4);
  B.charge_and_spin();
  std::vector<std::vector<double>> E;
  for (auto& b : B.blocks) E.push_back(std::vector<double>(b.size(), 0.0));
  return new Site(make_site(B, E, B.block_ops(B.fdag)));
}

Site* site_kondo(double J, double S) {
  const int d = (int)(2. * S + 1.);
  M Sz(d, d), Sp(d, d);
  for (int i = 0; i < d; i++) {
    const double mm = -S + i;
    Sz(i, i) = mm;
    if (i != d - 1) Sp(i, i + 1) = std::sqrt(S * (S + 1.) - mm * (mm + 1));
  }
  const M Sm = tr(Sp), Sx = scal(lin(Sp, 1, Sm, 1), 0.5), Sy = scal(lin(Sp, 1, Sm, -1), 0.5);
  Basis F;
  F.jordan_wigner(2);
  M cu = F.fdag[0], cd = F.fdag[1];
  const M wx = scal(lin(mul(cu, tr(cd)), 1, mul(cd, tr(cu)), 1), 0.5);
  const M wy = scal(lin(mul(cd, tr(cu)), 1, mul(cu, tr(cd)), -1), 0.5);
  const M wz = scal(lin(mul(cu, tr(cu)), 1, mul(cd, tr(cd)), -1), 0.5);
  const M wN = lin(mul(cu, tr(cu)), 1, mul(cd, tr(cd)), 1);
  const M H = scal(lin(lin(kron(Sz, wz), 1, kron(Sx, wx), 1), 1, kron(Sy, wy), -1), J);
  const M SzF = kron(Sz, eye(wx.r)), wzF = kron(eye(d), wz), wNF = kron(eye(d), wN);
  cu = kron(eye(d), cu);
  cd = kron(eye(d), cd);
  Basis B;
  const M I = eye(SzF.r);
  B.occ = {[&] { M t = scal(lin(I, 1, SzF, 2.), 0.5); std::vector<double> v(t.r); for (int i = 0; i < t.r; i++) v[i] = t(i, i); return v; }(),
           [&] { M t = scal(lin(I, 1, SzF, -2.), 0.5); std::vector<double> v(t.r); for (int i = 0; i < t.r; i++) v[i] = t(i, i); return v; }(),
           [&] { M t = scal(lin(wNF, 1, wzF, 2.), 0.5); std::vector<double> v(t.r); for (int i = 0; i < t.r; i++) v[i] = t(i, i); return v; }(),
           [&] { M t = scal(lin(wNF, 1, wzF, -2.), 0.5); std::vector<double> v(t.r); for (int i = 0; i < t.r; i++) v[i] = t(i, i); return v; }()};
  B.charge_and_spin();
  std::vector<std::vector<double>> E;
  std::vector<M> Ub;
  for (auto& hb : B.block_diag(H)) {
    M U;
    E.push_back(small_eigh(hb, U));
    Ub.push_back(U);
  }
  auto ops = B.block_ops({cu, cd});
  for (auto& op : ops)
    for (auto& [k, m] : op) m = mul(tr(Ub[k.first]), mul(m, Ub[k.second]));
  return new Site(make_site(B, E, ops));
}

Site* site_fdag_chi_spinhalf() {
  Basis B;
  B.jordan_wigner(2);
  B.charge_and_spin();
  const int D = 4;
  M chi(D, D);
  for (int s = 0; s < D; s++) chi(s, s) = (__builtin_popcount(s) & 1) ? -1.0 : 1.0;
  std::vector<M> ops;
  for (auto& f : B.fdag) ops.push_back(mul(f, chi));
  std::vector<std::vector<double>> E;
  for (auto& b : B.blocks) E.push_back(std::vector<double>(b.size(), 0.0));
  return new Site(make_site(B, E, B.block_ops(ops)));
}

// The double occupancy n_up n_dn of the spin-1/2 impurity in its sector blocks (the reference's
// spinhalf::doubleOccupancy, nrgcore/include/models/spinhalf.hpp:106-114): a system operator for
// the thermodynamic expectation values.
Site* site_double_occupancy_spinhalf() {
  Basis B;
  B.jordan_wigner(2);
  B.charge_and_spin();
  const M n_up = mul(B.fdag[0], tr(B.fdag[0])), n_dn = mul(B.fdag[1], tr(B.fdag[1]));
  std::vector<std::vector<double>> E;
  for (auto& b : B.blocks) E.push_back(std::vector<double>(b.size(), 0.0));
  return new Site(make_site(B, E, B.block_ops({mul(n_up, n_dn)})));
}

}  // namespace nrgk

struct nrgk_site {
  nrgk::Site s;
};

using nrgk::Error;

extern "C" {
int nrgk_site_spinhalf(double eps, double U, double h, nrgk_site** out) {
  if (!out) return NRGK_ERR_INVALID;
  nrgk::Site* s = nrgk::site_spinhalf(eps, U, h);
  *out = new nrgk_site{std::move(*s)};
  delete s;
  return NRGK_OK;
}
int nrgk_site_two_channel(nrgk_site** out) {
  if (!out) return NRGK_ERR_INVALID;
  nrgk::Site* s = nrgk::site_two_channel();
  *out = new nrgk_site{std::move(*s)};
  delete s;
  return NRGK_OK;
}
int nrgk_site_kondo(double J, double S, nrgk_site** out) {
  if (!out) return NRGK_ERR_INVALID;
  nrgk::Site* s = nrgk::site_kondo(J, S);
  *out = new nrgk_site{std::move(*s)};
  delete s;
  return NRGK_OK;
}
int nrgk_site_fdag_chi_spinhalf(nrgk_site** out) {
  if (!out) return NRGK_ERR_INVALID;
  nrgk::Site* s = nrgk::site_fdag_chi_spinhalf();
  *out = new nrgk_site{std::move(*s)};
  delete s;
  return NRGK_OK;
}
int nrgk_site_double_occupancy_spinhalf(nrgk_site** out) {
  if (!out) return NRGK_ERR_INVALID;
  nrgk::Site* s = nrgk::site_double_occupancy_spinhalf();
  *out = new nrgk_site{std::move(*s)};
  delete s;
  return NRGK_OK;
}
int nrgk_site_duplicate_ops(nrgk_site* site) {
  if (!site) return NRGK_ERR_INVALID;
  auto& s = site->s;
  if (s.n_ops < 2) return NRGK_ERR_INVALID;
  // append copies of operators 0 and 1 (examples/twoChannelSiam/main.cpp:24-25)
  size_t b0 = (size_t)s.nblocks[0], b1 = (size_t)s.nblocks[1];
  size_t d0 = 0, d01 = 0;
  for (size_t b = 0; b < b0; b++) d0 += (size_t)s.dims[s.bi[b]] * s.dims[s.bj[b]];
  d01 = d0;
  for (size_t b = b0; b < b0 + b1; b++) d01 += (size_t)s.dims[s.bi[b]] * s.dims[s.bj[b]];
  std::vector<int> bi(s.bi.begin(), s.bi.begin() + b0 + b1), bj(s.bj.begin(), s.bj.begin() + b0 + b1);
  std::vector<double> data(s.data.begin(), s.data.begin() + d01);
  // the new operators go at the end
  s.bi.insert(s.bi.end(), bi.begin(), bi.end());
  s.bj.insert(s.bj.end(), bj.begin(), bj.end());
  s.data.insert(s.data.end(), data.begin(), data.end());
  s.nblocks.push_back((int)b0);
  s.nblocks.push_back((int)b1);
  s.n_ops += 2;
  return NRGK_OK;
}
int nrgk_site_view(const nrgk_site* site, nrgk_site_desc* out) {
  if (!site || !out) return NRGK_ERR_INVALID;
  *out = site->s.view();
  return NRGK_OK;
}
int nrgk_site_destroy(nrgk_site* site) {
  delete site;
  return NRGK_OK;
}
}
