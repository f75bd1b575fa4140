// Observables, FDM spectrum and thermodynamics on the device.
#pragma once
#include <memory>
#include <string>

#include "chain.hpp"

namespace nrgk {

// System operators rotated along the chain (update_system_operator, sysOperator.hpp:35-133).
struct Obs {
  Chain* ch = nullptr;
  int nops = 0;
  std::vector<OpMap> blocks;  // full n_i x n_j blocks in the current basis
  DevBuf buf[2];
  int cur = 0;
  DevBuf dseg, dC, dY;
  Obs(Chain* c, int n_ops, const int* nblocks, const int* bi, const int* bj, const double* data);
  void update();
};

// One saved iteration (what NrgData::saveCurrentData + saveqOperator write, nrgData.hpp:134-205).
struct Shell {
  int cnt = 0;
  std::vector<int> n, eoff;
  std::vector<long long> uoff;
  std::vector<std::vector<std::pair<int, int>>> coupled;
  std::vector<std::vector<int>> sub_kidx;
  std::vector<int> pre_kept;  // eigenvaluesQ_kept_indices sizes at save time (previous iteration)
  std::vector<std::vector<int>> qn, pre_qn;  // current_sysmQ, pre_sysmQ
  int n_bath = 0;                            // bath sectors (coupled_nQ_index = i * n_bath + j)
  std::vector<double> E;      // eigenvalues (pre-shift), concatenated
  // diagonal of the (i, i) blocks of the system operators, [op][state] as E (0 where no block);
  // empty when the shell was recorded without a system operator
  std::vector<double> Q;
  int q_ops = 0;
  // U blocks, eigenvalues and observable blocks of the shell: sub-allocations of the recorder's arena
  struct Ref {
    void* p = nullptr;
    template <class T>
    T* as() const {
      return static_cast<T*>(p);
    }
  } dU, dE, dO;
  std::vector<OpMap> ops;
};

struct Fdm {
  Chain* ch = nullptr;
  Obs* obs = nullptr;
  Obs* sysop = nullptr;  // fdmThermodynamics::setSystemOperator
  int nbins = 0;
  double e_min = 1e-10, e_max = 10, delE = 0;
  std::vector<std::unique_ptr<Shell>> shells;
  // per operator, the signed weights w+ + w- of the last backward pass summed over its shells
  std::vector<double> signed_tot;
  // shells live in large chunks, filled in order: a record allocates nothing once a chunk has room
  std::vector<std::unique_ptr<DevBuf>> chunks;
  size_t chunk_used = 0;
  void* arena_alloc(size_t bytes);
  DevBuf dpos, dneg, dsum, dbf, dred[2], dRA, dAR, dT1, drho, ditems, dpt, dE, dth, dQ;
  Fdm(Chain* c, Obs* o, int nb, double emin, double emax);
  void record();
  // NrgData::saveCurrentData + saveqOperator (nrgData.hpp:134-205) of every recorded shell
  void save(const std::string& path, const std::string& op_group) const;
  // NrgData::loadCurrentData + loadqOperator (nrgData.hpp:214-293) of every shell of a saved file
  void load(const std::string& path, const std::string& op_group);
  void backward(double lambda, int scale_offset, int first_cnt, double* pos, double* neg, double* traces,
                double* sums, int cap, int* n_shells);
  void backward_shells(const std::vector<std::pair<Shell*, double>>& list, double* pos, double* neg, double* traces,
                       double* sums, int cap, int* n_shells);
  void thermo_shells(const std::vector<std::pair<Shell*, double>>& list, int nrg_max, const double* T, int nT,
                     double* S, double* C, double* ops_out);
  Shell* shell_at(int cnt) const;
  // shells registered one at a time (calcSpectrum / calcThermodynamics of the reference's driver loop):
  // (iteration counter, energy scale) in call order
  std::vector<std::pair<int, double>> registered;
  int loaded_nrg_max = 0;  // set by load(): the forward pass's final iteration count
  // ops_out ([sysop ops][nT], may be null): the system operators' thermal expectation values
  void thermo(double lambda, int scale_offset, int first_cnt, const double* T, int nT, double* S, double* C,
              double* ops_out = nullptr);
};

}  // namespace nrgk
