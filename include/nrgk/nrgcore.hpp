// C++ host API of libnrgk with the names and semantics of the reference's header-only nrgcore
// (nrgcore/include/nrgcore/nrgcore.hpp, sysOperator.hpp, dynamics/fdmSpectrum.hpp,
// static/fdmThermodynamics.hpp), layered on the C ABI of include/nrgk.h.  The numerical state of a
// chain lives on the GPU; the public data members that reference callers read (all_eigenvalue,
// eigenvaluesQ, ...) are host mirrors refreshed after every call that changes them.
//
// A driver of the reference compiles against it with its includes and model types changed: the
// nrgcore members it reads and writes (all_eigenvalue, eigenvaluesQ, pre_sysmQ, coupled_nQ_index,
// current_hamiltonQ[s], bath_eigenvaluesQ, nq_bath, a writable int nrg_iterations_cnt, ...),
// getPreWilsonSiteOperators(), update_system_operator, NrgData (saveCurrentData / saveqOperator /
// loadCurrentData / loadqOperator), fdmSpectrum (setOperator, calcSpectrum(scale), saveFinalData),
// fdmThermodynamics (the (nrg*, temperatures) constructor, calcThermodynamics(scale), saveFinalData)
// and h5stream::h5stream (write, close).  Differences a caller of the reference sees, all stated:
//   * model types are nrgk::site (spinhalf, twoChannel, spinhalfKondo, or any basis given as
//     arrays); nrgcore<im_type, bath_type> takes any type with a desc() returning nrgk_site_desc;
//   * current_hamiltonQ[s] returns a device copy of the block when it is read (the blocks stay on
//     the device); the other members are host mirrors refreshed after every call that changes them;
//   * NrgData keeps the shells on the device (not in a temporary HDF5 file): saveCurrentData records
//     the shell together with the system operators created on the chain, so saveqOperator and
//     loadqOperator have nothing left to do, and loadCurrentData selects the shell at the chain's
//     nrg_iterations_cnt without overwriting the host mirrors;
//   * calcSpectrum(scale) / calcThermodynamics(scale) register the loaded shell and its scale, and
//     saveFinalData runs the backward pass over the registered shells in one go;
//   * errors are std::runtime_error carrying nrgk_last_error(), as the reference throws
//     std::runtime_error (nrgcore.hpp:117-152).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../nrgk.h"

namespace nrgk {

inline void check(int rc) {
  if (rc != NRGK_OK) throw std::runtime_error(std::string("nrgk: ") + nrgk_last_error());
}

// Dense row-major block: the layout of the reference's qmatrix<> (qmatrix.hpp:66-75).
struct block {
  size_t rows = 0, cols = 0;
  std::vector<double> data;
  double& at(size_t i, size_t j) { return data[i * cols + j]; }
  double at(size_t i, size_t j) const { return data[i * cols + j]; }
  size_t getrow() const { return rows; }
  size_t getcol() const { return cols; }
};

// Block-sparse operator {(i, j) -> block}: the reference's qOperator (qOperator.hpp:18-158).
struct q_operator {
  std::map<std::pair<size_t, size_t>, block> blocks;
  std::optional<const block*> get(size_t i, size_t j) const {
    auto it = blocks.find({i, j});
    if (it == blocks.end()) return std::nullopt;
    return &it->second;
  }
  void set(size_t i, size_t j, block b) {
    if (!blocks.emplace(std::make_pair(i, j), std::move(b)).second) throw std::runtime_error("Found operator already");
  }
};

// Wilson-chain hopping t_n (examples/rgflowSIAM/main.cpp:5-10).
inline double hopping(size_t n, double lambda) {
  const double x = static_cast<double>(n);
  return 0.5 * (1.0 + 1.0 / lambda) * (1.0 - std::pow(lambda, -x - 1.0)) /
         std::sqrt((1.0 - std::pow(lambda, -2.0 * x - 1.0)) * (1.0 - std::pow(lambda, -2.0 * x - 3.0)));
}

// A site model in the reference's model concept (nrgcore.hpp:46-48,68-76).
class site {
 public:
  static site spinhalf(double eps, double U, double h = 0.0) {
    nrgk_site* s = nullptr;
    check(nrgk_site_spinhalf(eps, U, h, &s));
    return site(s);
  }
  static site twoChannel() {
    nrgk_site* s = nullptr;
    check(nrgk_site_two_channel(&s));
    return site(s);
  }
  static site spinhalfKondo(double J, double S = 0.5) {
    nrgk_site* s = nullptr;
    check(nrgk_site_kondo(J, S, &s));
    return site(s);
  }
  // f^dagger_sigma chi on the spinhalf impurity (examples/fdmSpectrumSiam/main.cpp:40-60)
  static site fdagChiSpinhalf() {
    nrgk_site* s = nullptr;
    check(nrgk_site_fdag_chi_spinhalf(&s));
    return site(s);
  }
  // spinhalf::getDoubleOcc (nrgcore/include/models/spinhalf.hpp:106-114, 215): n_up n_dn
  static site doubleOccupancySpinhalf() {
    nrgk_site* s = nullptr;
    check(nrgk_site_double_occupancy_spinhalf(&s));
    return site(s);
  }
  // spinless(eps), spinnlessTwoLead(eps), resonantTwoLead(eps, V_L, V_R, U), spinhalfSC with a pairing
  // term, and a free Wilson site in any symmetry (nrgcore/include/models/*.hpp)
  static site spinless(double eps = 0.0) {
    nrgk_site* s = nullptr;
    check(nrgk_site_spinless(eps, &s));
    return site(s);
  }
  static site spinlessTwoLead(double eps = 0.0) {
    nrgk_site* s = nullptr;
    check(nrgk_site_spinless_two_lead(eps, &s));
    return site(s);
  }
  static site resonantTwoLead(double eps, double v_left, double v_right, double U, bool observables = false) {
    nrgk_site* s = nullptr;
    check(nrgk_site_resonant_two_lead(eps, v_left, v_right, U, observables ? 1 : 0, &s));
    return site(s);
  }
  static site spinhalfSC(double eps, double U, double h = 0.0, double delta = 0.0) {
    nrgk_site* s = nullptr;
    check(nrgk_site_spinhalf_sc(eps, U, h, delta, &s));
    return site(s);
  }
  static site freeSite(int symmetry = 0) {
    nrgk_site* s = nullptr;
    check(nrgk_site_free(symmetry, &s));
    return site(s);
  }
  // any model: per-state quantum numbers, a Hamiltonian and full operators (row-major)
  static site fromHamiltonian(int n_states, int n_qn, const std::vector<int>& qn_state, const std::vector<double>& H,
                              const std::vector<std::vector<double>>& ops) {
    std::vector<double> flat;
    for (auto& o : ops) flat.insert(flat.end(), o.begin(), o.end());
    nrgk_site* s = nullptr;
    check(nrgk_site_from_hamiltonian(n_states, n_qn, qn_state.data(), H.data(), (int)ops.size(), flat.data(), &s));
    return site(s);
  }
  // the impurity f^dagger list doubled for two channels (examples/twoChannelSiam/main.cpp:24-25)
  site& duplicate_ops() {
    check(nrgk_site_duplicate_ops(h_.get()));
    return *this;
  }
  nrgk_site_desc desc() const {
    nrgk_site_desc d;
    check(nrgk_site_view(h_.get(), &d));
    return d;
  }
  std::vector<std::vector<int>> get_basis() const {
    const nrgk_site_desc d = desc();
    std::vector<std::vector<int>> q(d.n_sectors);
    for (int s = 0; s < d.n_sectors; s++) q[s].assign(d.qn + s * d.n_qn, d.qn + (s + 1) * d.n_qn);
    return q;
  }
  std::vector<std::vector<double>> get_eigenvaluesQ() const {
    const nrgk_site_desc d = desc();
    std::vector<std::vector<double>> e(d.n_sectors);
    size_t o = 0;
    for (int s = 0; s < d.n_sectors; s++) {
      e[s].assign(d.energies + o, d.energies + o + d.dims[s]);
      o += d.dims[s];
    }
    return e;
  }
  std::vector<double> get_chi_Q() const {
    const nrgk_site_desc d = desc();
    return std::vector<double>(d.chi, d.chi + d.n_sectors);
  }
  size_t n_ops() const { return static_cast<size_t>(desc().n_ops); }

 private:
  explicit site(nrgk_site* s) : h_(s, [](nrgk_site* p) { nrgk_site_destroy(p); }) {}
  std::shared_ptr<nrgk_site> h_;
};

// A device context (GPU ordinal + HBM admission capacity).
class context {
 public:
  explicit context(int device = 0) {
    nrgk_ctx* c = nullptr;
    check(nrgk_ctx_create(device, &c));
    h_.reset(c, [](nrgk_ctx* p) { nrgk_ctx_destroy(p); });
  }
  static context& default_context() {
    static context c(0);
    return c;
  }
  void set_capacity(size_t bytes) { check(nrgk_ctx_set_capacity(h_.get(), bytes)); }
  nrgk_ctx* get() const { return h_.get(); }

 private:
  std::shared_ptr<nrgk_ctx> h_;
};

// What NrgData, fdmSpectrum and fdmThermodynamics need of a chain, whatever its model types: its
// handle, the iteration counter the drivers decrement, the system operators created on it and the
// recorder of its shells (created by the chain's NrgData).
class chain_base {
 public:
  virtual ~chain_base() = default;
  nrgk_chain* handle() const { return h_.get(); }
  // nrg_iterations_cnt / nrg_iterations_min (nrgcore.hpp:536-538); drivers decrement the counter in
  // the backward loop (examples/fdmSpectrumSiam/main.cpp:163), the next call that reads it sees that
  int nrg_iterations_cnt = 0;
  int nrg_iterations_min = 0;
  // pushes a counter the caller has changed into the chain
  void sync_counter() {
    int c = 0, m = 0;
    check(nrgk_chain_counters(h_.get(), &c, &m));
    if (c != nrg_iterations_cnt) check(nrgk_chain_set_counter(h_.get(), nrg_iterations_cnt));
  }
  void register_operator(nrgk_obs* o) { sysops_.push_back(o); }
  nrgk_obs* first_operator() const { return sysops_.empty() ? nullptr : sysops_.front(); }
  std::shared_ptr<nrgk_fdm> recorder;

 protected:
  std::shared_ptr<nrgk_chain> h_;
  std::vector<nrgk_obs*> sysops_;
};

// nrgcore<im_type, bath_type> (nrgcore.hpp:25-550).  The models must outlive the chain only as
// long as it is being constructed: their descriptions are copied (the reference keeps raw pointers
// and re-reads the bath on every add_bath_site, nrgcore.hpp:68-70; set_bath() gives that here).
template <class im_type = site, class bath_type = site>
class nrgcore : public chain_base {
 public:
  nrgcore(im_type& impurity, bath_type& bath, context& ctx = context::default_context(),
          int mode = NRGK_MODE_MIXED, size_t max_kept = 1024)
      : current_hamiltonQ(this) {
    const nrgk_site_desc di = impurity.desc(), db = bath.desc();
    nrgk_chain* c = nullptr;
    check(nrgk_chain_create(ctx.get(), &di, &db, max_kept, mode, &c));
    h_.reset(c, [](nrgk_chain* p) { nrgk_chain_destroy(p); });
    n_ops_ = static_cast<size_t>(di.n_ops);
    nqn_ = di.n_qn;
    set_bath_mirrors(db);
  }
  nrgcore(const nrgcore&) = delete;
  nrgcore& operator=(const nrgcore&) = delete;
  // set_parameters (nrgcore.hpp:336-339)
  void set_parameters(size_t n = 1024) {
    check(nrgk_chain_set_parameters(h_.get(), n));
    refresh_counters();
  }
  void set_bath(bath_type& bath) {
    const nrgk_site_desc db = bath.desc();
    check(nrgk_chain_set_bath(h_.get(), &db));
    set_bath_mirrors(db);
  }
  // add_bath_site (nrgcore.hpp:65-97): assembly, batched eigensolve, f^dagger rotation
  void add_bath_site(const std::vector<double>& thopping, double rescale) {
    sync_counter();
    check(nrgk_chain_add_bath_site(h_.get(), thopping.data(), static_cast<int>(thopping.size()), rescale));
    refresh();
  }
  // update_internal_state (nrgcore.hpp:105-109): truncation, shift, counters
  void update_internal_state() {
    sync_counter();
    check(nrgk_chain_update_internal_state(h_.get()));
    refresh();
  }
  std::vector<std::vector<int>> get_basis_nQ() {
    if (current_sysmQ.empty()) throw std::runtime_error("current_sysmQ.empty!");
    return current_sysmQ;
  }
  std::vector<std::vector<double>> get_eigenvaluesQ() {
    if (eigenvaluesQ.empty()) throw std::runtime_error("eigenvaluesQ.empty!");
    return eigenvaluesQ;
  }
  // the rotated Wilson-site f^dagger of the current iteration (nrgcore.hpp:377), kept x kept
  std::vector<q_operator> get_f_dag_operator() {
    std::vector<q_operator> out(n_ops_);
    for (size_t p = 0; p < n_ops_; p++) {
      int nb = 0;
      check(nrgk_chain_fdag_blocks(h_.get(), (int)p, &nb, nullptr, nullptr, nullptr, nullptr, nullptr, 0));
      std::vector<int> bi(nb), bj(nb), r(nb), c(nb);
      check(nrgk_chain_fdag_blocks(h_.get(), (int)p, &nb, bi.data(), bj.data(), r.data(), c.data(), nullptr, 0));
      size_t tot = 0;
      for (int b = 0; b < nb; b++) tot += (size_t)r[b] * c[b];
      std::vector<double> data(tot);
      check(nrgk_chain_fdag_blocks(h_.get(), (int)p, &nb, bi.data(), bj.data(), r.data(), c.data(), data.data(),
                                   data.size()));
      size_t o = 0;
      for (int b = 0; b < nb; b++) {
        block blk{(size_t)r[b], (size_t)c[b], {}};
        blk.data.assign(data.begin() + o, data.begin() + o + (size_t)r[b] * c[b]);
        o += blk.data.size();
        out[p].set((size_t)bi[b], (size_t)bj[b], std::move(blk));
      }
    }
    return out;
  }
  // checkHigherEnergyDiscarded (nrgcore.hpp:384-386)
  bool checkHigherEnergyDiscarded() {
    int v = 0;
    check(nrgk_chain_higher_energy_discarded(h_.get(), &v));
    return v != 0;
  }
  // getPreWilsonSiteOperators (nrgcore.hpp:548): the rotated f-dagger of the last site in the kept
  // basis, the operators the next add_bath_site couples to (kept x kept blocks)
  std::vector<q_operator>* getPreWilsonSiteOperators() {
    pre_wilson_ = get_f_dag_operator();
    return &pre_wilson_;
  }
  // the reference drivers decrement nrg_iterations_cnt during the backward pass
  void set_nrg_iterations_cnt(int c) {
    nrg_iterations_cnt = c;
    sync_counter();
  }
  // re-reads the host mirrors after the chain was stepped through the C ABI (step_chains)
  void refresh_host_mirrors() { refresh(); }

  // current_hamiltonQ (nrgcore.hpp:526): the U block of a sector (eigenvectors in its columns,
  // dsyevd's layout, qmatrix.hpp:390-391), copied from the device when read
  class u_blocks {
   public:
    explicit u_blocks(nrgcore* c) : c_(c) {}
    block operator[](size_t sector) const {
      const size_t n = c_->eigenvaluesQ.at(sector).size();
      block b{n, n, std::vector<double>(n * n)};
      check(nrgk_chain_block_u(c_->h_.get(), (int)sector, b.data.data(), b.data.size()));
      return b;
    }
    block operator()(size_t sector) const { return (*this)[sector]; }
    size_t size() const { return c_->eigenvaluesQ.size(); }

   private:
    nrgcore* c_;
  };

  // public data of the reference (nrgcore.hpp:520-542), host mirrors
  std::vector<double> all_eigenvalue;
  std::vector<double> relativeGroundStateEnergy;
  std::vector<std::vector<size_t>> eigenvaluesQ_kept_indices;
  u_blocks current_hamiltonQ;
  std::vector<std::vector<int>> current_sysmQ;
  std::vector<std::vector<int>> pre_sysmQ;
  std::vector<std::vector<double>> eigenvaluesQ;
  std::vector<std::vector<size_t>> coupled_nQ_index;
  std::vector<std::vector<double>> bath_eigenvaluesQ;
  std::vector<std::vector<int>> nq_bath;

 private:
  void set_bath_mirrors(const nrgk_site_desc& db) {
    nq_bath.assign(db.n_sectors, std::vector<int>(db.n_qn));
    bath_eigenvaluesQ.assign(db.n_sectors, {});
    size_t o = 0;
    for (int s = 0; s < db.n_sectors; s++) {
      for (int q = 0; q < db.n_qn; q++) nq_bath[s][q] = db.qn[s * db.n_qn + q];
      bath_eigenvaluesQ[s].assign(db.energies + o, db.energies + o + db.dims[s]);
      o += db.dims[s];
    }
  }
  void refresh_counters() {
    check(nrgk_chain_counters(h_.get(), &nrg_iterations_cnt, &nrg_iterations_min));
  }
  void refresh() {
    refresh_counters();
    size_t n = 0;
    check(nrgk_chain_all_eigenvalues(h_.get(), nullptr, 0, &n));
    all_eigenvalue.resize(n);
    check(nrgk_chain_all_eigenvalues(h_.get(), all_eigenvalue.data(), n, &n));
    check(nrgk_chain_relative_gs(h_.get(), nullptr, 0, &n));
    relativeGroundStateEnergy.resize(n);
    check(nrgk_chain_relative_gs(h_.get(), relativeGroundStateEnergy.data(), n, &n));
    int S = 0;
    check(nrgk_chain_sectors(h_.get(), &S, nullptr, nullptr));
    std::vector<int> qn((size_t)S * nqn_), dims(S);
    check(nrgk_chain_sectors(h_.get(), &S, qn.data(), dims.data()));
    current_sysmQ.assign(S, std::vector<int>(nqn_));
    for (int s = 0; s < S; s++)
      for (int q = 0; q < nqn_; q++) current_sysmQ[s][q] = qn[(size_t)s * nqn_ + q];
    size_t total = 0;
    check(nrgk_chain_eigenvalues_q(h_.get(), nullptr, 0, &total));
    std::vector<double> e(total);
    check(nrgk_chain_eigenvalues_q(h_.get(), e.data(), e.size(), &total));
    eigenvaluesQ.assign(S, {});
    size_t o = 0;
    for (int s = 0; s < S; s++) {
      eigenvaluesQ[s].assign(e.begin() + o, e.begin() + o + dims[s]);
      o += dims[s];
    }
    int Sk = 0;
    check(nrgk_chain_kept_counts(h_.get(), nullptr, 0, &Sk));
    std::vector<int> kept(Sk);
    check(nrgk_chain_kept_counts(h_.get(), kept.data(), Sk, &Sk));
    eigenvaluesQ_kept_indices.assign(Sk, {});
    for (int s = 0; s < Sk; s++)
      for (int k = 0; k < kept[s]; k++) eigenvaluesQ_kept_indices[s].push_back((size_t)k);
    int Sp = 0;
    check(nrgk_chain_pre_sectors(h_.get(), &Sp, nullptr));
    std::vector<int> pq((size_t)Sp * nqn_);
    check(nrgk_chain_pre_sectors(h_.get(), &Sp, pq.data()));
    pre_sysmQ.assign(Sp, std::vector<int>(nqn_));
    for (int s = 0; s < Sp; s++)
      for (int q = 0; q < nqn_; q++) pre_sysmQ[s][q] = pq[(size_t)s * nqn_ + q];
    size_t nc = 0;
    check(nrgk_chain_coupled_index(h_.get(), nullptr, nullptr, 0, &nc));
    std::vector<int> counts(S), flat(nc);
    check(nrgk_chain_coupled_index(h_.get(), counts.data(), flat.data(), flat.size(), &nc));
    coupled_nQ_index.assign(S, {});
    size_t co = 0;
    for (int s = 0; s < S; s++)
      for (int k = 0; k < counts[s]; k++) coupled_nQ_index[s].push_back((size_t)flat[co++]);
  }
  size_t n_ops_ = 0;
  int nqn_ = 2;
  std::vector<q_operator> pre_wilson_;
};

// System operators rotated along the chain: the std::vector<qOperator> argument of the reference's
// update_system_operator (sysOperator.hpp:35-133), resident on the device.
class system_operator {
 public:
  system_operator(chain_base& chain, const site& op) {
    const nrgk_site_desc d = op.desc();
    nrgk_obs* o = nullptr;
    check(nrgk_obs_create(chain.handle(), d.n_ops, d.op_nblocks, d.op_bi, d.op_bj, d.op_data, &o));
    h_.reset(o, [](nrgk_obs* p) { nrgk_obs_destroy(p); });
    n_ops_ = (size_t)d.n_ops;
    chain.register_operator(o);
  }
  nrgk_obs* handle() const { return h_.get(); }
  size_t size() const { return n_ops_; }

 private:
  std::shared_ptr<nrgk_obs> h_;
  size_t n_ops_ = 0;
};

// update_system_operator(nrg, ops) (sysOperator.hpp:35-37): after add_bath_site, before
// update_internal_state; the operators move to the new eigenbasis in place.
template <class chain_type>
void update_system_operator(chain_type* /*nrg*/, system_operator& ops) {
  check(nrgk_obs_update(ops.handle()));
}

// h5stream::h5stream (nrgcore/include/utils/h5stream.hpp:244-478): an HDF5 file written dataset by
// dataset; a vector of vectors becomes the datasets name0, name1, ... (h5stream.hpp:295-305).
namespace h5stream {
class h5stream {
 public:
  explicit h5stream(const std::string& path) : path_(path) { check(nrgk_h5_create(path.c_str())); }
  void write(const std::vector<double>& v, const std::string& name) {
    check(nrgk_h5_write(path_.c_str(), name.c_str(), v.data(), v.size()));
  }
  void write(const std::vector<std::vector<double>>& v, const std::string& name) {
    for (size_t i = 0; i < v.size(); i++) write(v[i], name + std::to_string(i));
  }
  void close() {}
  const std::string& path() const { return path_; }

 private:
  std::string path_;
};
}  // namespace h5stream

// NrgData (nrgcore/include/nrgcore/nrgData.hpp:31-302) on the device: saveCurrentData records the
// iteration (U blocks, eigenvalues, sectors, coupled index, kept counts -- what the reference writes to
// its temporary HDF5 file) together with the system operators created on the chain; loadCurrentData
// selects the shell at the chain's nrg_iterations_cnt for the calcSpectrum / calcThermodynamics
// that follow.  The first chain operator, if any, is the recorded observable (and the system operator
// of the thermal expectation values).
class NrgData {
 public:
  explicit NrgData(chain_base* nrg, int n_bins = 100000) : nrg_(nrg), n_bins_(n_bins) {}
  // saveCurrentData (nrgData.hpp:134-164): after add_bath_site and update_system_operator
  void saveCurrentData() {
    if (!nrg_->recorder) {
      nrgk_fdm* f = nullptr;
      nrgk_obs* op = nrg_->first_operator();
      check(nrgk_fdm_create(nrg_->handle(), op, n_bins_, 1e-10, 10.0, &f));
      nrg_->recorder.reset(f, [](nrgk_fdm* p) { nrgk_fdm_destroy(p); });
      if (op) check(nrgk_fdm_set_system_operator(f, op));
    }
    nrg_->sync_counter();
    check(nrgk_fdm_record(nrg_->recorder.get()));
  }
  // saveqOperator (nrgData.hpp:174-205): the operators went into the shell with saveCurrentData
  template <class op_type>
  void saveqOperator(op_type*, const std::string& = "") {}
  // loadCurrentData (nrgData.hpp:266-293): the shell at the (decremented) nrg_iterations_cnt
  void loadCurrentData() {
    if (!nrg_->recorder) throw std::runtime_error("NrgData: nothing was saved");
    nrg_->sync_counter();
  }
  template <class op_type>
  void loadqOperator(op_type*, const std::string& = "") {}
  // NrgData::saveCurrentData's HDF5 layout of every recorded shell, written to a file on request
  void save(const std::string& path, const std::string& op_group = "/nparticle") {
    if (!nrg_->recorder) throw std::runtime_error("NrgData: nothing was saved");
    check(nrgk_fdm_save(nrg_->recorder.get(), path.c_str(), op_group.empty() ? nullptr : op_group.c_str()));
  }
  bool debugIO = false;

 private:
  chain_base* nrg_;
  int n_bins_;
};

// fdmSpectrum (dynamics/fdmSpectrum.hpp:24-364), T = 0.  Two forms: the reference's (constructed
// on the chain's address, setOperator, calcSpectrum(scale) per loaded shell, saveFinalData into an
// h5stream), and a record-forward / backward-once form (constructed with the chain and the operator,
// record() where the drivers save the iteration, backward() over all shells).
class fdmSpectrum {
 public:
  fdmSpectrum(chain_base& chain, system_operator& ops, int n_bins = 100000, double e_min = 1e-10,
              double e_max = 10.0)
      : n_bins_(n_bins), n_ops_(ops.size()) {
    nrgk_fdm* f = nullptr;
    check(nrgk_fdm_create(chain.handle(), ops.handle(), n_bins, e_min, e_max, &f));
    h_.reset(f, [](nrgk_fdm* p) { nrgk_fdm_destroy(p); });
  }
  // the reference's form (fdmSpectrum.hpp:35): the shells come from the chain's NrgData
  explicit fdmSpectrum(chain_base* nrg) : nrg_(nrg), n_bins_(100000) {}
  // setOperator (fdmSpectrum.hpp:306-314): the recorded operators; a separate a-operator is not
  // supported (the reference indexes an empty matrix with one, fdmSpectrum.hpp:221-224)
  void setOperator(system_operator* b, system_operator* a = nullptr) {
    if (a && a != b) throw std::runtime_error("fdmSpectrum::setOperator: a separate a-operator is not supported");
    if (nrg_ && nrg_->first_operator() != b->handle())
      throw std::runtime_error("fdmSpectrum::setOperator: not the operator the chain's NrgData recorded");
    n_ops_ = b->size();
  }
  // calcSpectrum (fdmSpectrum.hpp:55-63) for the shell NrgData::loadCurrentData selected
  void calcSpectrum(double energyScale) {
    if (!nrg_ || !nrg_->recorder) throw std::runtime_error("fdmSpectrum::calcSpectrum: no recorded shells");
    nrg_->sync_counter();
    check(nrgk_fdm_add_shell(nrg_->recorder.get(), nrg_->nrg_iterations_cnt, energyScale));
  }
  // saveFinalData (fdmSpectrum.hpp:323-337): the backward pass over the shells calcSpectrum took, then
  // GreenFnEnergyPoints, GreenFnPositiveWeight{p}, GreenFnNegativeWeight{p}
  void saveFinalData(h5stream::h5stream* pfile) {
    if (!nrg_ || !nrg_->recorder) throw std::runtime_error("fdmSpectrum::saveFinalData: no recorded shells");
    nrgk_fdm* f = nrg_->recorder.get();
    std::vector<double> pos(n_ops_ * n_bins_), neg(n_ops_ * n_bins_), tr(4096), sums(4096);
    int ns = 0;
    check(nrgk_fdm_run_spectrum(f, pos.data(), neg.data(), tr.data(), sums.data(), (int)tr.size(), &ns));
    positiveWeight.assign(n_ops_, {});
    negativeWeight.assign(n_ops_, {});
    for (size_t p = 0; p < n_ops_; p++) {
      positiveWeight[p].assign(pos.begin() + p * n_bins_, pos.begin() + (p + 1) * n_bins_);
      negativeWeight[p].assign(neg.begin() + p * n_bins_, neg.begin() + (p + 1) * n_bins_);
    }
    rhoTraces.assign(tr.begin(), tr.begin() + ns);
    std::vector<double> e(n_bins_);
    check(nrgk_fdm_energy_points(f, e.data()));
    pfile->write(e, "GreenFnEnergyPoints");
    pfile->write(positiveWeight, "GreenFnPositiveWeight");
    pfile->write(negativeWeight, "GreenFnNegativeWeight");
  }
  // where the reference drivers save the iteration (examples/fdmSpectrumSiam/main.cpp:90-93)
  void record() { check(nrgk_fdm_record(h_.get())); }
  // the backward loop of calcSpectrum(Lambda^{-(in-1)/2}) calls (examples/fdmSpectrumSiam/main.cpp:110-122)
  void backward(double lambda, int scale_offset = -1, int first_cnt = 0) {
    positiveWeight.assign(n_ops_, std::vector<double>(n_bins_, 0.0));
    negativeWeight.assign(n_ops_, std::vector<double>(n_bins_, 0.0));
    std::vector<double> pos(n_ops_ * n_bins_), neg(n_ops_ * n_bins_), tr(4096), sums(4096);
    int ns = 0;
    check(nrgk_fdm_backward(h_.get(), lambda, scale_offset, first_cnt, pos.data(), neg.data(), tr.data(), sums.data(),
                            (int)tr.size(), &ns));
    for (size_t p = 0; p < n_ops_; p++) {
      positiveWeight[p].assign(pos.begin() + p * n_bins_, pos.begin() + (p + 1) * n_bins_);
      negativeWeight[p].assign(neg.begin() + p * n_bins_, neg.begin() + (p + 1) * n_bins_);
    }
    rhoTraces.assign(tr.begin(), tr.begin() + ns);
  }
  std::vector<double> energyPoints() const {
    std::vector<double> e(n_bins_);
    check(nrgk_fdm_energy_points(h_.get(), e.data()));
    return e;
  }
  nrgk_fdm* handle() const { return h_.get(); }
  // NrgData::saveCurrentData + saveqOperator of every recorded shell, in the reference's HDF5 layout
  void save(const std::string& path, const std::string& op_group = "/nparticle") {
    check(nrgk_fdm_save(h_.get(), path.c_str(), op_group.empty() ? nullptr : op_group.c_str()));
  }

  std::vector<std::vector<double>> positiveWeight, negativeWeight;
  std::vector<double> rhoTraces;

 private:
  std::shared_ptr<nrgk_fdm> h_;
  chain_base* nrg_ = nullptr;
  size_t n_bins_, n_ops_ = 0;
};

// Wilson chain of the z-shifted logarithmic discretisation of a bath (independFiles/zAverage:
// calcTn, hoppingElement.cpp:16-105, for the flat band; zAverage.hpp:12-195 for a density tabulated
// at x): {t_n Lambda^{n/2}, eps_n Lambda^{n/2}}, n < n_sites.  quad 0: exact integrals of the
// tabulated density, 1: the reference's 100-point left sums.
inline std::pair<std::vector<double>, std::vector<double>> z_chain(double z, double lambda, int n_sites,
                                                                   double D = 1.0, const std::vector<double>& x = {},
                                                                   const std::vector<double>& rho = {}, int quad = 0,
                                                                   int n_intervals = 0, bool normalize = true) {
  if (x.size() != rho.size()) throw std::runtime_error("z_chain: x and rho differ in length");
  std::vector<double> t(n_sites), e(n_sites);
  check(nrgk_zdisc_chain(z, lambda, D, x.data(), rho.data(), (int)x.size(), quad, n_intervals, normalize ? 1 : 0,
                         n_sites, t.data(), e.data()));
  return {t, e};
}

// qmatrix::diag (nrgcore/include/utils/qmatrix.hpp:375-404) for many blocks in one call
// (nrgk_eigh_batched): each block's data is replaced by its eigenvectors (in the columns, row-major,
// as dsyevd leaves A) and its ascending eigenvalues are returned.
inline std::vector<std::vector<double>> diag_blocks(std::vector<block>& blocks, context& ctx = context::default_context(),
                                                    int mode = NRGK_MODE_MIXED) {
  std::vector<int> dims;
  std::vector<double> H;
  for (const auto& b : blocks) {
    if (b.rows != b.cols || b.data.size() != b.rows * b.cols) throw std::runtime_error("diag_blocks: square blocks only");
    dims.push_back((int)b.rows);
    H.insert(H.end(), b.data.begin(), b.data.end());
  }
  size_t ne = 0;
  for (int n : dims) ne += (size_t)n;
  std::vector<double> w(ne), U(H.size());
  check(nrgk_eigh_batched(ctx.get(), (int)dims.size(), dims.data(), H.data(), w.data(), U.data(), mode));
  std::vector<std::vector<double>> out;
  size_t eo = 0, uo = 0;
  for (auto& b : blocks) {
    const size_t n = b.rows;
    out.emplace_back(w.begin() + eo, w.begin() + eo + n);
    std::copy(U.begin() + uo, U.begin() + uo + n * n, b.data.begin());
    eo += n;
    uo += n * n;
  }
  return out;
}

// One Wilson site on each of several independent chains, run concurrently (nrgk_chains_step; the
// reference steps one chain at a time).  Chain i does what the reference's FDM drivers do per site
// (examples/fdmSpectrumSiam/main.cpp:86-95): add_bath_site(t[i], rescale[i]); update_system_operator
// on ops[i] when ops is given (one per chain); record into recorders[i] (the handle() of an
// fdmSpectrum or fdmThermodynamics of that chain, or null) when higher energies were discarded;
// update_internal_state.  Returns which chains recorded; the chains' host mirrors are refreshed.
template <class chain_type, class op_type = system_operator>
inline std::vector<bool> step_chains(const std::vector<chain_type*>& chains, const std::vector<std::vector<double>>& t,
                                     const std::vector<double>& rescale, const std::vector<op_type*>& ops = {},
                                     const std::vector<nrgk_fdm*>& recorders = {}, int n_threads = 0) {
  const size_t n = chains.size();
  if (t.size() != n || rescale.size() != n || (!ops.empty() && ops.size() != n) ||
      (!recorders.empty() && recorders.size() != n))
    throw std::runtime_error("step_chains: one entry per chain in t, rescale, ops and recorders");
  const size_t nt = n ? t[0].size() : 0;
  std::vector<double> tf;
  for (const auto& v : t) {
    if (v.size() != nt) throw std::runtime_error("step_chains: all chains take the same number of hoppings");
    tf.insert(tf.end(), v.begin(), v.end());
  }
  std::vector<nrgk_chain*> hc(n);
  std::vector<nrgk_obs*> ho;
  for (size_t i = 0; i < n; i++) {
    hc[i] = chains[i]->handle();
    if (!ops.empty()) ho.push_back(ops[i]->handle());
  }
  std::vector<int> rec(n, 0);
  check(nrgk_chains_step(hc.data(), (int)n, tf.data(), (int)nt, rescale.data(), ho.data(), ops.empty() ? 0 : 1,
                         recorders.empty() ? nullptr : recorders.data(), n_threads, rec.data()));
  for (auto* c : chains) c->refresh_host_mirrors();
  return std::vector<bool>(rec.begin(), rec.end());
}

// fdmThermodynamics (static/fdmThermodynamics.hpp:17-260): shells recorded during the forward pass,
// entropy and specific heat on a temperature grid.
class fdmThermodynamics {
 public:
  explicit fdmThermodynamics(chain_base& chain) {
    nrgk_fdm* f = nullptr;
    check(nrgk_fdm_create(chain.handle(), nullptr, 1, 1e-10, 10.0, &f));
    h_.reset(f, [](nrgk_fdm* p) { nrgk_fdm_destroy(p); });
  }
  // the reference's form (fdmThermodynamics.hpp:26-37): constructed after the forward pass on the
  // chain's address and the temperatures; the shells come from the chain's NrgData
  fdmThermodynamics(chain_base* nrg, std::vector<double> tArray)
      : temperatureArray(std::move(tArray)), nrg_(nrg), nrg_max_(nrg->nrg_iterations_cnt) {}
  // setSystemOperator (fdmThermodynamics.hpp:39-41), reference form: the operator NrgData recorded
  void setSystemOperator(system_operator* ops) {
    if (!nrg_ || nrg_->first_operator() != ops->handle())
      throw std::runtime_error("fdmThermodynamics::setSystemOperator: not the operator the chain's NrgData recorded");
    n_sysop_ = ops->size();
  }
  // calcThermodynamics (fdmThermodynamics.hpp:42-89) for the shell NrgData::loadCurrentData selected
  void calcThermodynamics(double energyRescale) {
    if (!nrg_ || !nrg_->recorder) throw std::runtime_error("fdmThermodynamics::calcThermodynamics: no recorded shells");
    nrg_->sync_counter();
    check(nrgk_fdm_add_shell(nrg_->recorder.get(), nrg_->nrg_iterations_cnt, energyRescale));
  }
  // saveFinalData (fdmThermodynamics.hpp:101-236): entropy and specific heat over the shells
  // calcThermodynamics took; temperatureArray, entropy, specificHeat (and qsysOPeratorExpValue{q})
  void saveFinalData(h5stream::h5stream* pfile) {
    if (!nrg_ || !nrg_->recorder) throw std::runtime_error("fdmThermodynamics::saveFinalData: no recorded shells");
    const size_t nT = temperatureArray.size();
    entropy.assign(nT, 0.0);
    specificHeat.assign(nT, 0.0);
    std::vector<double> v(std::max<size_t>(1, n_sysop_ * nT));
    check(nrgk_thermo_run(nrg_->recorder.get(), nrg_max_, temperatureArray.data(), (int)nT, entropy.data(),
                          specificHeat.data(), n_sysop_ ? v.data() : nullptr));
    qsysOPeratorExpValue.clear();
    for (size_t q = 0; q < n_sysop_; q++) qsysOPeratorExpValue.emplace_back(v.begin() + q * nT, v.begin() + (q + 1) * nT);
    pfile->write(temperatureArray, "temperatureArray");
    pfile->write(entropy, "entropy");
    pfile->write(specificHeat, "specificHeat");
    if (n_sysop_) pfile->write(qsysOPeratorExpValue, "qsysOPeratorExpValue");
  }
  void record() { check(nrgk_fdm_record(h_.get())); }
  // setSystemOperator (fdmThermodynamics.hpp:39-41): shells recorded from now on keep the diagonals of
  // the operators, which the caller updates (update_system_operator) before each record
  void setSystemOperator(system_operator& ops) {
    check(nrgk_fdm_set_system_operator(h_.get(), ops.handle()));
    n_sysop_ = ops.size();
  }
  void calcThermodynamics(const std::vector<double>& T, double lambda, int scale_offset = -1, int first_cnt = 0) {
    temperatureArray = T;
    entropy.assign(T.size(), 0.0);
    specificHeat.assign(T.size(), 0.0);
    check(nrgk_thermo(h_.get(), lambda, scale_offset, first_cnt, T.data(), (int)T.size(), entropy.data(),
                      specificHeat.data()));
    qsysOPeratorExpValue.clear();
    if (n_sysop_) {
      std::vector<double> v(n_sysop_ * T.size());
      check(nrgk_thermo_expectation(h_.get(), lambda, scale_offset, first_cnt, T.data(), (int)T.size(), v.data()));
      for (size_t q = 0; q < n_sysop_; q++)
        qsysOPeratorExpValue.emplace_back(v.begin() + q * T.size(), v.begin() + (q + 1) * T.size());
    }
  }
  std::vector<double> temperatureArray, entropy, specificHeat;
  std::vector<std::vector<double>> qsysOPeratorExpValue;  // [op][T] (fdmThermodynamics.hpp:148-150, 206)

 private:
  std::shared_ptr<nrgk_fdm> h_;
  chain_base* nrg_ = nullptr;
  int nrg_max_ = 0;
  size_t n_sysop_ = 0;
};

}  // namespace nrgk
