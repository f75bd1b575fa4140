// C entry points of the FP64 oracle, for the Python test harness (ctypes).  Test infrastructure
// only: see nrg_oracle.hpp.  No exception crosses this boundary; failures return -1 and the
// message is kept for orc_last_error().
#include <cstring>
#include <memory>
#include <string>

#include "nrg_oracle.hpp"

using namespace nrg_oracle;

namespace {
thread_local std::string g_err;
template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
struct ChainH {
  std::unique_ptr<Chain> c;
};
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
const char* orc_lapack_library() {
  try {
    return Lapack::get().library.c_str();
  } catch (const std::exception& e) {
    g_err = e.what();
    return "";
  }
}
void orc_set_threads(int n) {
  auto& L = Lapack::get();
  if (L.set_threads) L.set_threads(n);
}

// ---------------------------------------------------------------- models
void* orc_model_spinhalf(double eps, double U, double h) {
  SiteModel* m = nullptr;
  guard([&] { m = new SiteModel(make_spinhalf(eps, U, h)); });
  return m;
}
void* orc_model_two_channel() {
  SiteModel* m = nullptr;
  guard([&] { m = new SiteModel(make_two_channel()); });
  return m;
}
void* orc_model_kondo(double J, double S) {
  SiteModel* m = nullptr;
  guard([&] { m = new SiteModel(make_kondo(J, S)); });
  return m;
}
// append copies of f[0], f[1] (examples/twoChannelSiam/main.cpp:24-25)
void orc_model_dup_ops(void* mp) {
  auto* m = static_cast<SiteModel*>(mp);
  auto a = m->f_dag_operator[0], b = m->f_dag_operator[1];
  m->f_dag_operator.push_back(a);
  m->f_dag_operator.push_back(b);
}
void orc_model_free(void* m) { delete static_cast<SiteModel*>(m); }
int orc_model_n_sectors(void* mp) { return (int)static_cast<SiteModel*>(mp)->n_Q.size(); }
int orc_model_n_ops(void* mp) { return (int)static_cast<SiteModel*>(mp)->f_dag_operator.size(); }
// qn[s*2+q], dims[s], energies (concatenated), chi[s]
void orc_model_sectors(void* mp, int* qn, int* dims, double* energies, double* chi) {
  auto* m = static_cast<SiteModel*>(mp);
  size_t off = 0;
  for (size_t s = 0; s < m->n_Q.size(); s++) {
    for (size_t q = 0; q < m->n_Q[s].size(); q++) qn[s * m->n_Q[s].size() + q] = m->n_Q[s][q];
    dims[s] = (int)m->eigenvalues_Q[s].size();
    for (double e : m->eigenvalues_Q[s]) energies[off++] = e;
    chi[s] = m->chi_Q[s];
  }
}
int orc_model_op_nblocks(void* mp, int p) { return (int)static_cast<SiteModel*>(mp)->f_dag_operator[p].blocks.size(); }
// bi, bj per block, data concatenated row-major
void orc_model_op_blocks(void* mp, int p, int* bi, int* bj, double* data) {
  auto* m = static_cast<SiteModel*>(mp);
  size_t b = 0, off = 0;
  for (auto& [k, mat] : m->f_dag_operator[p].blocks) {
    bi[b] = (int)k.first;
    bj[b] = (int)k.second;
    std::memcpy(data + off, mat.a.data(), mat.a.size() * sizeof(double));
    off += mat.a.size();
    b++;
  }
}

// ---------------------------------------------------------------- chain
void* orc_chain_new(void* imp, void* bath, long K) {
  ChainH* h = nullptr;
  guard([&] {
    h = new ChainH;
    h->c = std::make_unique<Chain>(*static_cast<SiteModel*>(imp), *static_cast<SiteModel*>(bath));
    h->c->set_parameters((size_t)K);
  });
  return h;
}
void orc_chain_free(void* h) { delete static_cast<ChainH*>(h); }
void orc_chain_keep_h(void* h, int on) { static_cast<ChainH*>(h)->c->keep_hamiltonians = on != 0; }
int orc_add_bath_site(void* h, const double* t, int nt, double rescale) {
  return guard([&] { static_cast<ChainH*>(h)->c->add_bath_site(std::vector<double>(t, t + nt), rescale); });
}
int orc_update_internal_state(void* h) {
  return guard([&] { static_cast<ChainH*>(h)->c->update_internal_state(); });
}
int orc_cnt(void* h) { return static_cast<ChainH*>(h)->c->nrg_iterations_cnt; }
int orc_syevd_failures(void* h) { return static_cast<ChainH*>(h)->c->syevd_failures; }
void orc_step_times(void* h, double* out4) {
  auto& t = static_cast<ChainH*>(h)->c->times;
  out4[0] = t.assemble;
  out4[1] = t.diag;
  out4[2] = t.rotate;
  out4[3] = t.truncate;
}
long orc_n_all_eigenvalues(void* h) { return (long)static_cast<ChainH*>(h)->c->all_eigenvalue.size(); }
void orc_all_eigenvalues(void* h, double* out) {
  auto& v = static_cast<ChainH*>(h)->c->all_eigenvalue;
  std::memcpy(out, v.data(), v.size() * sizeof(double));
}
long orc_n_relgs(void* h) { return (long)static_cast<ChainH*>(h)->c->relativeGroundStateEnergy.size(); }
void orc_relgs(void* h, double* out) {
  auto& v = static_cast<ChainH*>(h)->c->relativeGroundStateEnergy;
  std::memcpy(out, v.data(), v.size() * sizeof(double));
}
// sectors of eigenvaluesQ (current after add_bath_site, previous-kept view after update)
int orc_n_sectors(void* h) { return (int)static_cast<ChainH*>(h)->c->eigenvaluesQ.size(); }
void orc_sector_dims(void* h, int* dims) {
  auto& c = *static_cast<ChainH*>(h)->c;
  for (size_t s = 0; s < c.eigenvaluesQ.size(); s++) dims[s] = (int)c.eigenvaluesQ[s].size();
}
void orc_eigenvalues_q(void* h, double* out) {
  auto& c = *static_cast<ChainH*>(h)->c;
  size_t off = 0;
  for (auto& v : c.eigenvaluesQ)
    for (double e : v) out[off++] = e;
}
int orc_n_kept_sectors(void* h) { return (int)static_cast<ChainH*>(h)->c->eigenvaluesQ_kept_indices.size(); }
void orc_kept_counts(void* h, int* out) {
  auto& c = *static_cast<ChainH*>(h)->c;
  for (size_t s = 0; s < c.eigenvaluesQ_kept_indices.size(); s++) out[s] = (int)c.eigenvaluesQ_kept_indices[s].size();
}
int orc_n_qn(void* h) {
  auto& c = *static_cast<ChainH*>(h)->c;
  return c.pre_sysmQ.empty() ? 0 : (int)c.pre_sysmQ[0].size();
}
int orc_n_pre_sectors(void* h) { return (int)static_cast<ChainH*>(h)->c->pre_sysmQ.size(); }
void orc_pre_qn(void* h, int* out) {
  auto& c = *static_cast<ChainH*>(h)->c;
  size_t k = 0;
  for (auto& q : c.pre_sysmQ)
    for (int x : q) out[k++] = x;
}
int orc_n_cur_sectors(void* h) { return (int)static_cast<ChainH*>(h)->c->current_sysmQ.size(); }
void orc_cur_qn(void* h, int* out) {
  auto& c = *static_cast<ChainH*>(h)->c;
  size_t k = 0;
  for (auto& q : c.current_sysmQ)
    for (int x : q) out[k++] = x;
}
// H block s before the eigensolve (requires orc_chain_keep_h)
void orc_block_h(void* h, int s, double* out) {
  auto& m = static_cast<ChainH*>(h)->c->hamiltonians_before_diag[s];
  std::memcpy(out, m.a.data(), m.a.size() * sizeof(double));
}
void orc_block_u(void* h, int s, double* out) {
  auto& m = static_cast<ChainH*>(h)->c->current_hamiltonQ[s];
  std::memcpy(out, m.a.data(), m.a.size() * sizeof(double));
}
int orc_n_ops(void* h) { return (int)static_cast<ChainH*>(h)->c->pre_fdag.size(); }
int orc_fdag_nblocks(void* h, int p) { return (int)static_cast<ChainH*>(h)->c->pre_fdag[p].blocks.size(); }
void orc_fdag_block_info(void* h, int p, int* bi, int* bj, int* rows, int* cols) {
  size_t b = 0;
  for (auto& [k, m] : static_cast<ChainH*>(h)->c->pre_fdag[p].blocks) {
    bi[b] = (int)k.first;
    bj[b] = (int)k.second;
    rows[b] = (int)m.r;
    cols[b] = (int)m.c;
    b++;
  }
}
void orc_fdag_block(void* h, int p, int bidx, double* out) {
  auto it = static_cast<ChainH*>(h)->c->pre_fdag[p].blocks.begin();
  std::advance(it, bidx);
  std::memcpy(out, it->second.a.data(), it->second.a.size() * sizeof(double));
}

// ---------------------------------------------------------------- complete driver runs
// One Wilson-chain driver in the form of the reference's examples:
//   first_V > 0 : add_bath_site({V..}, 1.0) first (examples/rgflowSIAM/main.cpp:29), then sites
//                 in = 0..nsites-1 with rescale = in > 0 ? sqrt(L) : 1 (main.cpp:33-42);
//   first_V <= 0: sites in = 0..nsites-1 with rescale sqrt(L) (examples/entropyKondo/main.cpp:62-65).
struct OrcDriver {
  int imp_kind;  // 0 spinhalf(eps,U,h), 1 kondo(J,S)
  double eps, U, h, J, S;
  int bath_kind;  // 0 spinhalf(0,0), 1 twoChannel
  int dup_ops;    // duplicate the impurity f-dagger operators (two-channel)
  double Lambda;
  double first_V;
  int nsites;
  long K;
};

static SiteModel driver_imp(const OrcDriver& d) {
  SiteModel m = d.imp_kind == 1 ? make_kondo(d.J, d.S) : make_spinhalf(d.eps, d.U, d.h);
  if (d.dup_ops) {
    auto a = m.f_dag_operator[0], b = m.f_dag_operator[1];
    m.f_dag_operator.push_back(a);
    m.f_dag_operator.push_back(b);
  }
  return m;
}
static SiteModel driver_bath(const OrcDriver& d) {
  return d.bath_kind == 1 ? make_two_channel() : make_spinhalf(0, 0);
}

// RG flow: all_eigenvalue after every loop site, concatenated; counts[n] = its length.
int orc_run_rgflow(const OrcDriver* d, double* eig_out, long cap, long* counts) {
  return guard([&] {
    SiteModel imp = driver_imp(*d), bath = driver_bath(*d);
    Chain c(imp, bath);
    c.set_parameters((size_t)d->K);
    size_t nop = imp.f_dag_operator.size();
    if (d->first_V > 0) {
      c.add_bath_site(std::vector<double>(nop, d->first_V), 1.0);
      c.update_internal_state();
    }
    long off = 0;
    for (int in = 0; in < d->nsites; in++) {
      double rescale = d->first_V > 0 ? (in > 0 ? std::sqrt(d->Lambda) : 1.0) : std::sqrt(d->Lambda);
      c.add_bath_site(std::vector<double>(nop, hopping(in, d->Lambda)), rescale);
      c.update_internal_state();
      counts[in] = (long)c.all_eigenvalue.size();
      if (off + counts[in] > cap) throw std::runtime_error("orc_run_rgflow: output capacity");
      std::memcpy(eig_out + off, c.all_eigenvalue.data(), c.all_eigenvalue.size() * sizeof(double));
      off += counts[in];
    }
  });
}

// Thermodynamics driver (examples/entropySIAM/main.cpp:13-96, examples/entropyKondo/main.cpp:35-119).
// min_iter0 = the driver's initial minIterations (0 for SIAM, 1 for the Kondo/free drivers).
static int run_thermo(const OrcDriver* d, const double* hop, int min_iter0, const double* T, int nT, double* S_out,
                      double* C_out, int low_clamp_zero) {
  return guard([&] {
    SiteModel imp = driver_imp(*d), bath = driver_bath(*d);
    Chain c(imp, bath);
    c.set_parameters((size_t)d->K);
    size_t nop = imp.f_dag_operator.size();
    if (d->first_V > 0) {
      c.add_bath_site(std::vector<double>(nop, d->first_V), 1.0);
      c.update_internal_state();
    }
    std::map<int, Shell> shells;
    size_t minIterations = (size_t)min_iter0;
    for (int in = 0; in < d->nsites; in++) {
      double rescale = d->first_V > 0 ? (in > 0 ? std::sqrt(d->Lambda) : 1.0) : std::sqrt(d->Lambda);
      c.add_bath_site(std::vector<double>(nop, hop ? hop[in] : hopping(in, d->Lambda)), rescale);
      if (c.checkHigherEnergyDiscarded()) shells[c.nrg_iterations_cnt] = save_shell(c);
      else minIterations++;
      c.update_internal_state();
    }
    FdmThermo th(c.nrg_iterations_cnt, c.bath_eigenvaluesQ, std::vector<double>(T, T + nT));
    th.low_clamp_zero = low_clamp_zero != 0;
    for (size_t in = (size_t)d->nsites; in > minIterations; in--) {
      double sc = std::pow(d->Lambda, -(1.0 * (double)in - 1.0) / 2.0);
      c.nrg_iterations_cnt--;
      th.calc(shells.at(c.nrg_iterations_cnt), sc);
    }
    auto [S, C] = th.finish();
    std::memcpy(S_out, S.data(), S.size() * sizeof(double));
    std::memcpy(C_out, C.data(), C.size() * sizeof(double));
  });
}

int orc_run_thermo(const OrcDriver* d, int min_iter0, const double* T, int nT, double* S_out, double* C_out,
                   int low_clamp_zero) {
  return run_thermo(d, nullptr, min_iter0, T, nT, S_out, C_out, low_clamp_zero);
}
// the same driver with the chain's hoppings given per site (a z-shifted discretisation's t_n(z)
// in place of hopping(n, Lambda))
int orc_run_thermo_t(const OrcDriver* d, const double* hop, int min_iter0, const double* T, int nT, double* S_out,
                     double* C_out) {
  return run_thermo(d, hop, min_iter0, T, nT, S_out, C_out, 0);
}

// Thermodynamics driver with system operators on a spinhalf impurity (fdmThermodynamics::
// setSystemOperator, static/fdmThermodynamics.hpp:39-41, 63-87, 191-211): the operators are rotated
// with update_system_operator after every add_bath_site and saved beside each shell, then read back
// by calcThermodynamics in the backward loop of examples/entropySIAM/main.cpp:79-92.
// op_kind 0: {n_up n_dn} (spinhalf.hpp:114); 1: {n_up, n_dn, identity}.  V_out: [n_ops][nT].
int orc_run_thermo_ops(const OrcDriver* d, int min_iter0, const double* T, int nT, int op_kind, double* S_out,
                       double* C_out, double* V_out, int* n_ops) {
  return guard([&] {
    if (d->imp_kind != 0) throw std::runtime_error("orc_run_thermo_ops: spinhalf impurity only");
    SiteModel imp = driver_imp(*d), bath = driver_bath(*d);
    FermionBasis fb(2);
    const Mat nu = dot(fb.fdag_full[0], transpose(fb.fdag_full[0]));
    const Mat nd = dot(fb.fdag_full[1], transpose(fb.fdag_full[1]));
    std::vector<BlockOp> ops = op_kind == 0 ? imp.double_occupancy : fb.block_operators({nu, nd, eye(4)});
    *n_ops = (int)ops.size();
    Chain c(imp, bath);
    c.set_parameters((size_t)d->K);
    size_t nop = imp.f_dag_operator.size();
    std::map<int, Shell> shells;
    size_t minIterations = (size_t)min_iter0;
    if (d->first_V > 0) {
      c.add_bath_site(std::vector<double>(nop, d->first_V), 1.0);
      update_system_operator(c, ops);
      c.update_internal_state();
    }
    for (int in = 0; in < d->nsites; in++) {
      double rescale = d->first_V > 0 ? (in > 0 ? std::sqrt(d->Lambda) : 1.0) : std::sqrt(d->Lambda);
      c.add_bath_site(std::vector<double>(nop, hopping(in, d->Lambda)), rescale);
      update_system_operator(c, ops);
      if (c.checkHigherEnergyDiscarded()) {
        shells[c.nrg_iterations_cnt] = save_shell(c);
        shells[c.nrg_iterations_cnt].ops = ops;
      } else {
        minIterations++;
      }
      c.update_internal_state();
    }
    FdmThermo th(c.nrg_iterations_cnt, c.bath_eigenvaluesQ, std::vector<double>(T, T + nT));
    for (size_t in = (size_t)d->nsites; in > minIterations; in--) {
      double sc = std::pow(d->Lambda, -(1.0 * (double)in - 1.0) / 2.0);
      c.nrg_iterations_cnt--;
      const Shell& sh = shells.at(c.nrg_iterations_cnt);
      th.calc(sh, sc, &sh.ops);
    }
    auto [S, C] = th.finish();
    std::memcpy(S_out, S.data(), S.size() * sizeof(double));
    std::memcpy(C_out, C.data(), C.size() * sizeof(double));
    for (size_t q = 0; q < th.qsysOPeratorExpValue.size(); q++)
      std::memcpy(V_out + q * nT, th.qsysOPeratorExpValue[q].data(), nT * sizeof(double));
  });
}

// T = 0 FDM spectral function of f^dagger chi (examples/fdmSpectrumSiam/main.cpp:61-187, with the
// hopping list sized to the operator count).  pos/neg: [n_ops][nbins]; traces/sums per shell.
// planted mistake for the FDM backward pass (FdmSpectrum::plant), used by the oracle's own tests
static int g_fdm_plant = 0;
void orc_set_fdm_plant(int p) { g_fdm_plant = p; }

int orc_run_fdm_signed(const OrcDriver* d, int nbins, double* pos, double* neg, double* traces, double* sums,
                       int* n_shells, double* signed_tot);
int orc_run_fdm(const OrcDriver* d, int nbins, double* pos, double* neg, double* traces, double* sums,
                int* n_shells) {
  return orc_run_fdm_signed(d, nbins, pos, neg, traces, sums, n_shells, nullptr);
}
// orc_run_fdm, also returning per operator the signed weights summed over all shells
int orc_run_fdm_signed(const OrcDriver* d, int nbins, double* pos, double* neg, double* traces, double* sums,
                       int* n_shells, double* signed_tot) {
  return guard([&] {
    SiteModel imp = driver_imp(*d), bath = driver_bath(*d);
    Chain c(imp, bath);
    c.set_parameters((size_t)d->K);
    size_t nop = imp.f_dag_operator.size();
    auto ops = make_fdag_chi_spinhalf();
    std::map<int, Shell> shells;
    std::map<int, std::vector<BlockOp>> saved_ops;
    size_t minIterations = 0;
    c.add_bath_site(std::vector<double>(nop, d->first_V), 1.0);
    update_system_operator(c, ops);
    c.update_internal_state();
    for (int in = 0; in < d->nsites; in++) {
      double rescale = in > 0 ? std::sqrt(d->Lambda) : 1.0;
      c.add_bath_site(std::vector<double>(nop, hopping(in, d->Lambda)), rescale);
      update_system_operator(c, ops);
      if (c.checkHigherEnergyDiscarded()) {
        shells[c.nrg_iterations_cnt] = save_shell(c);
        saved_ops[in] = ops;
      } else {
        minIterations++;
      }
      c.update_internal_state();
    }
    FdmSpectrum f(ops.size(), nbins);
    f.plant = g_fdm_plant;
    int ns = 0;
    for (size_t in = (size_t)d->nsites; in > minIterations; in--) {
      c.nrg_iterations_cnt--;
      Shell& sh = shells.at(c.nrg_iterations_cnt);
      sh.ops = saved_ops.at((int)in - 1);
      f.calc(sh, c.bath_eigenvaluesQ, std::pow(d->Lambda, -((double)in - 1.0) / 2.0));
      sh.ops.clear();
      ns++;
    }
    for (size_t p = 0; p < ops.size(); p++) {
      std::memcpy(pos + p * nbins, f.positiveWeight[p].data(), nbins * sizeof(double));
      std::memcpy(neg + p * nbins, f.negativeWeight[p].data(), nbins * sizeof(double));
    }
    for (int i = 0; i < ns; i++) {
      traces[i] = f.rhoTraces[i];
      sums[i] = f.specSums[i];
    }
    if (signed_tot)
      for (size_t p = 0; p < ops.size(); p++) signed_tot[p] = f.signedWeight[p];
    *n_shells = ns;
  });
}

// CPU baseline of the RG-flow drivers (examples/rgflowSIAM/main.cpp:29-42, twoChannelSiam): first site
// {V..V} with rescale 1, then loop sites in with rescale in > 0 ? sqrt(Lambda) : 1; each step is
// add_bath_site + update_internal_state and the all_eigenvalue read, timed with std::chrono over loop
// sites [n_skip, n_skip + n_time).
int orc_time_rgflow(const OrcDriver* d, int n_skip, int n_time, double* seconds) {
  return guard([&] {
    SiteModel imp = driver_imp(*d), bath = driver_bath(*d);
    Chain c(imp, bath);
    c.set_parameters((size_t)d->K);
    size_t nop = imp.f_dag_operator.size();
    c.add_bath_site(std::vector<double>(nop, d->first_V), 1.0);
    c.update_internal_state();
    double tot = 0, guard_sum = 0;
    for (int in = 0; in < n_skip + n_time; in++) {
      auto t0 = std::chrono::steady_clock::now();
      c.add_bath_site(std::vector<double>(nop, hopping(in, d->Lambda)), in > 0 ? std::sqrt(d->Lambda) : 1.0);
      c.update_internal_state();
      guard_sum += c.all_eigenvalue.empty() ? 0.0 : c.all_eigenvalue[0];
      auto t1 = std::chrono::steady_clock::now();
      if (in >= n_skip) tot += std::chrono::duration<double>(t1 - t0).count();
    }
    *seconds = guard_sum == guard_sum ? tot : -1.0;
  });
}

// CPU baseline: the FDM-driver forward pass in the reference's form (add_bath_site,
// update_system_operator of f^dagger chi, shell save, update_internal_state), timed over loop sites
// [n_skip, n_skip + n_time) with std::chrono like the reference's timer (nrgcore.hpp:66,90-93).
// Seconds spent in loop sites [n_skip, n_skip + n_time) at the current thread count and, when
// n_time1 > 0, in the following n_time1 sites at one thread (seconds1).
int orc_time_forward2(const OrcDriver* d, int n_skip, int n_time, int n_time1, double* seconds, double* seconds1);
int orc_time_forward(const OrcDriver* d, int n_skip, int n_time, double* seconds) {
  return orc_time_forward2(d, n_skip, n_time, 0, seconds, nullptr);
}
int orc_time_forward2(const OrcDriver* d, int n_skip, int n_time, int n_time1, double* seconds, double* seconds1) {
  return guard([&] {
    SiteModel imp = driver_imp(*d), bath = driver_bath(*d);
    Chain c(imp, bath);
    c.set_parameters((size_t)d->K);
    size_t nop = imp.f_dag_operator.size();
    auto ops = make_fdag_chi_spinhalf();
    std::vector<Shell> shells;
    c.add_bath_site(std::vector<double>(nop, d->first_V), 1.0);
    update_system_operator(c, ops);
    c.update_internal_state();
    double tot = 0, tot1 = 0;
    for (int in = 0; in < n_skip + n_time + n_time1; in++) {
      if (in == n_skip + n_time) orc_set_threads(1);
      auto t0 = std::chrono::steady_clock::now();
      double rescale = in > 0 ? std::sqrt(d->Lambda) : 1.0;
      c.add_bath_site(std::vector<double>(nop, hopping(in, d->Lambda)), rescale);
      update_system_operator(c, ops);
      if (c.checkHigherEnergyDiscarded()) {
        shells.push_back(save_shell(c));
        shells.back().ops = ops;
      }
      c.update_internal_state();
      auto t1 = std::chrono::steady_clock::now();
      const double dt = std::chrono::duration<double>(t1 - t0).count();
      if (in >= n_skip + n_time) tot1 += dt;
      else if (in >= n_skip) tot += dt;
    }
    *seconds = tot;
    if (seconds1) *seconds1 = tot1;
  });
}

void orc_broaden(const double* x, const double* w, int n, const double* omega, int no, double b, double* out) {
  auto r = broaden(std::vector<double>(x, x + n), std::vector<double>(w, w + n), std::vector<double>(omega, omega + no), b);
  std::memcpy(out, r.data(), r.size() * sizeof(double));
}
double orc_hopping(int n, double L) { return hopping(n, L); }

}  // extern "C"
