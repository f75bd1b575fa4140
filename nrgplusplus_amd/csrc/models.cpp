// Site models (host-side setup, not the hot path).
//
// Restates the reference's model construction:
//   fermionBasis (nrgcore/include/models/fermionBasis.hpp:20-442): Jordan-Wigner
//     f_i^dagger = I^(dof-i-1) (x) f^dagger (x) sigma_z^i, so orbital i is bit i of the Fock index
//     and <s|f_i^dagger|s'> = (-1)^(number of occupied orbitals below i);
//     chargeAndSpin quantum numbers [N(even orbitals), N(odd orbitals)] (:122-141);
//     sectors in first-appearance order (:230-268); a block is kept iff sum |elem| > 0 (:291-318).
//   chargeOnly [N] (:103-117) and spinOnly [N(even) - N(odd)] (:59-72, 150-193) quantum numbers;
//   spinhalf (nrgcore/include/models/spinhalf.hpp:81-124), twoChannel
//   (nrgcore/src/models/twoChannel.cpp:12-48), spinhalfKondo (nrgcore/include/models/spinhalfKondo.hpp:28-166),
//   spinless, spinnlessTwoLead, resonantTwoLead, spinhalfSC and freeModel (nrgcore/include/models/),
//   the f^dagger chi observable of examples/fdmSpectrumSiam/main.cpp:86-106, and a generic builder
//   from a Hamiltonian on any basis with per-state quantum numbers.
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
    group();
  }
  // chargeOnly: [N] (fermionBasis.hpp:103-117)
  void charge_only() {
    qn_state.clear();
    for (size_t s = 0; s < occ[0].size(); s++) {
      int n = 0;
      for (size_t j = 0; j < occ.size(); j++) n += (int)occ[j][s];
      qn_state.push_back({n});
    }
    group();
  }
  // spinOnly: [N(even orbitals) - N(odd orbitals)] (fermionBasis.hpp:59-72, 150-193)
  void spin_only() {
    qn_state.clear();
    for (size_t s = 0; s < occ[0].size(); s++) {
      int sz = 0;
      for (size_t j = 0; j < occ.size(); j++) sz += (j % 2 == 0 ? 1 : -1) * (int)occ[j][s];
      qn_state.push_back({sz});
    }
    group();
  }
  void symmetry(int sym) {
    if (sym == 0) charge_and_spin();
    else if (sym == 1) charge_only();
    else spin_only();
  }
  // sectors in first-appearance order, states ascending within a sector (fermionBasis.hpp:230-268)
  void group() {
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
  s.n_qn = B.unique_qn.empty() ? 0 : (int)B.unique_qn[0].size();
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

// the sectors, energies and parities of s with other operator blocks
static Site make_site_qn(const Site& s0, const std::vector<std::map<std::pair<int, int>, M>>& ops) {
  Site s = s0;
  s.nblocks.clear();
  s.bi.clear();
  s.bj.clear();
  s.data.clear();
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

// A site from a Hamiltonian on any basis whose states carry quantum numbers: states grouped into
// sectors (first appearance), each sector block diagonalised ascending, the operators rotated into
// the eigenbasis, U_i^T F U_j, where their unrotated block is nonzero -- the construction every model
// constructor of the reference repeats (e.g. resonantTwoLead.hpp:31-62, spinhalfSC.hpp:53-86).
static Site site_from_states(Basis& B, const M& H, const std::vector<M>& ops) {
  const int D = (int)B.qn_state.size();
  std::vector<int> sec(D);
  for (size_t i = 0; i < B.blocks.size(); i++)
    for (int st : B.blocks[i]) sec[st] = (int)i;
  for (int i = 0; i < D; i++)
    for (int j = 0; j < D; j++)
      if (sec[i] != sec[j] && H(i, j) != 0.0)
        throw Error(NRGK_ERR_INVALID, "nrgk: the Hamiltonian couples different quantum-number sectors");
  std::vector<std::vector<double>> E;
  std::vector<M> Ub;
  for (auto& hb : B.block_diag(H)) {
    M U;
    E.push_back(small_eigh(hb, U));
    Ub.push_back(U);
  }
  auto bops = B.block_ops(ops);
  for (auto& op : bops)
    for (auto& [k, m] : op) m = mul(tr(Ub[k.first]), mul(m, Ub[k.second]));
  return make_site(B, E, bops);
}

static M number(const M& fd) { return mul(fd, tr(fd)); }

// spinless(eps) (nrgcore/include/models/spinnless.hpp): one orbital, chargeOnly, H = eps n
Site* site_spinless(double eps) {
  Basis B;
  B.jordan_wigner(1);
  B.charge_only();
  return new Site(site_from_states(B, scal(number(B.fdag[0]), eps), B.fdag));
}

// spinnlessTwoLead(eps) (nrgcore/include/models/spinlessTwoLead.hpp): two orbitals, chargeOnly,
// H = eps n_0, both f^dagger coupling to the chain
Site* site_spinless_two_lead(double eps) {
  Basis B;
  B.jordan_wigner(2);
  B.charge_only();
  return new Site(site_from_states(B, scal(number(B.fdag[0]), eps), B.fdag));
}

// resonantTwoLead(eps, V_L, V_R, U) (nrgcore/include/models/resonantTwoLead.hpp:20-62): a level
// (orbital 0) and the first site of each lead (orbitals 1, 2), chargeOnly;
// H = eps n_0 + V_L (f0+ f1 + f1+ f0) + V_R (f0+ f2 + f2+ f0) + U (n_0 - I)(n_1 + n_2 - I), the
// identity as there (its `half_id` is n_label.id(), qmatrix.hpp:251-262).  The chain couples to the
// two lead orbitals;
// with observables = 1 the operators are instead the level's f^dagger and n_0 = f0+ f0 blocks
// (impurityFdag, impurityNparticle) for update_system_operator.
Site* site_resonant_two_lead(double eps, double vl, double vr, double U, bool observables) {
  Basis B;
  B.jordan_wigner(3);
  B.charge_only();
  const M& f0 = B.fdag[0];
  const M& f1 = B.fdag[1];
  const M& f2 = B.fdag[2];
  const M n0 = number(f0), n1 = number(f1), n2 = number(f2);
  const M I = eye(n0.r);
  M H = scal(n0, eps);
  H = lin(H, 1.0, lin(mul(f0, tr(f1)), 1.0, mul(f1, tr(f0)), 1.0), vl);
  H = lin(H, 1.0, lin(mul(f0, tr(f2)), 1.0, mul(f2, tr(f0)), 1.0), vr);
  H = lin(H, 1.0, mul(lin(n0, 1.0, I, -1.0), lin(lin(n1, 1.0, n2, 1.0), 1.0, I, -1.0)), U);
  if (!observables) return new Site(site_from_states(B, H, {f1, f2}));
  // impurityNparticle: F F^T of the rotated level operator, on the diagonal blocks (resonantTwoLead.hpp:64-74)
  Site s = site_from_states(B, H, {f0});
  std::vector<std::map<std::pair<int, int>, M>> ops(2);
  size_t off = 0;
  std::map<std::pair<int, int>, M> fd;
  for (int b = 0; b < s.nblocks[0]; b++) {
    const int i = s.bi[b], j = s.bj[b];
    M m(s.dims[i], s.dims[j]);
    std::copy(s.data.begin() + off, s.data.begin() + off + m.a.size(), m.a.begin());
    off += m.a.size();
    fd[{i, j}] = m;
  }
  ops[0] = fd;
  for (auto& [k, m] : fd) {
    M nn = mul(m, tr(m));
    auto& slot = ops[1][{k.first, k.first}];
    slot = slot.r ? lin(slot, 1.0, nn, 1.0) : nn;
  }
  return new Site(make_site_qn(s, ops));
}

// spinhalfSC(eps, U, h) with a pairing term (nrgcore/include/models/spinhalfSC.hpp:27-123): two
// orbitals, spinOnly; H = eps (n_up + n_dn) + h (n_up - n_dn) + U n_up n_dn + delta (f_up f_dn + h.c.).
// The reference's addSCDelta (:89-123) replaces H by the pairing term alone; the same site follows
// here with eps = h = U = 0.
Site* site_spinhalf_sc(double eps, double U, double h, double delta) {
  Basis B;
  B.jordan_wigner(2);
  B.spin_only();
  const M nu = number(B.fdag[0]), nd = number(B.fdag[1]);
  M H = lin(lin(lin(nu, 1.0, nd, 1.0), eps, lin(nu, 1.0, nd, -1.0), h), 1.0, mul(nu, nd), U);
  const M pair = lin(mul(tr(B.fdag[0]), tr(B.fdag[1])), 1.0, mul(B.fdag[1], B.fdag[0]), 1.0);
  H = lin(H, 1.0, pair, delta);
  return new Site(site_from_states(B, H, B.fdag));
}

// a free Wilson site in any of the three symmetries (freeModel.hpp for chargeAndSpin; the spinOnly
// site is the bath of the superconducting impurities, the chargeOnly one spinnlessTwoLead(0))
Site* site_free(int symmetry) {
  Basis B;
  B.jordan_wigner(2);
  B.symmetry(symmetry);
  return new Site(site_from_states(B, M(4, 4), B.fdag));
}

// a site from user data: D states with n_qn quantum numbers each, H (D x D) and n_ops operators
Site* site_from_hamiltonian(int D, int n_qn, const int* qn_state, const double* H, int n_ops, const double* ops) {
  Basis B;
  B.qn_state.clear();
  for (int i = 0; i < D; i++) B.qn_state.push_back(std::vector<int>(qn_state + (size_t)i * n_qn, qn_state + (size_t)(i + 1) * n_qn));
  B.group();
  M Hm(D, D);
  std::copy(H, H + (size_t)D * D, Hm.a.begin());
  std::vector<M> om;
  for (int p = 0; p < n_ops; p++) {
    M o(D, D);
    std::copy(ops + (size_t)p * D * D, ops + (size_t)(p + 1) * D * D, o.a.begin());
    om.push_back(o);
  }
  for (int i = 0; i < D; i++)
    for (int j = 0; j < i; j++)
      if (Hm(i, j) != Hm(j, i)) throw Error(NRGK_ERR_INVALID, "nrgk: the Hamiltonian is not symmetric");
  Site s = site_from_states(B, Hm, om);
  s.n_qn = n_qn;
  return new Site(s);
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

template <class F>
static int wrap(F&& make, nrgk_site** out) {
  try {
    nrgk::Site* s = make();
    *out = new nrgk_site{std::move(*s)};
    delete s;
    return NRGK_OK;
  } catch (const nrgk::Error& e) {
    nrgk::set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    nrgk::set_last_error(e.what());
    return NRGK_ERR_INVALID;
  }
}

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
int nrgk_site_spinless(double eps, nrgk_site** out) {
  if (!out) return NRGK_ERR_INVALID;
  return wrap([&] { return nrgk::site_spinless(eps); }, out);
}
int nrgk_site_spinless_two_lead(double eps, nrgk_site** out) {
  if (!out) return NRGK_ERR_INVALID;
  return wrap([&] { return nrgk::site_spinless_two_lead(eps); }, out);
}
int nrgk_site_resonant_two_lead(double eps, double v_left, double v_right, double U, int observables,
                                nrgk_site** out) {
  if (!out) return NRGK_ERR_INVALID;
  return wrap([&] { return nrgk::site_resonant_two_lead(eps, v_left, v_right, U, observables != 0); }, out);
}
int nrgk_site_spinhalf_sc(double eps, double U, double h, double delta, nrgk_site** out) {
  if (!out) return NRGK_ERR_INVALID;
  return wrap([&] { return nrgk::site_spinhalf_sc(eps, U, h, delta); }, out);
}
int nrgk_site_free(int symmetry, nrgk_site** out) {
  if (!out || symmetry < 0 || symmetry > 2) return NRGK_ERR_INVALID;
  return wrap([&] { return nrgk::site_free(symmetry); }, out);
}
int nrgk_site_from_hamiltonian(int n_states, int n_qn, const int* qn_state, const double* H, int n_ops,
                               const double* ops, nrgk_site** out) {
  if (!out || n_states <= 0 || n_qn <= 0 || !qn_state || !H || n_ops < 0 || (n_ops > 0 && !ops))
    return NRGK_ERR_INVALID;
  return wrap([&] { return nrgk::site_from_hamiltonian(n_states, n_qn, qn_state, H, n_ops, ops); }, out);
}
int nrgk_fock_basis(int n_orb, int symmetry, double* fdag, int* qn_state) {
  if (n_orb <= 0 || n_orb > 12 || symmetry < 0 || symmetry > 2) return NRGK_ERR_INVALID;
  nrgk::Basis B;
  B.jordan_wigner(n_orb);
  B.symmetry(symmetry);
  const size_t D = (size_t)1 << n_orb;
  if (fdag)
    for (int i = 0; i < n_orb; i++) std::copy(B.fdag[i].a.begin(), B.fdag[i].a.end(), fdag + (size_t)i * D * D);
  if (qn_state)
    for (size_t st = 0; st < D; st++)
      std::copy(B.qn_state[st].begin(), B.qn_state[st].end(), qn_state + st * B.qn_state[st].size());
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
