// Host-side chain state of libnrgk (what nrgcore<imp,bath> holds, nrgcore.hpp:512-550), with the
// numerical state resident on the device.
#pragma once
#include <algorithm>
#include <map>
#include <stdexcept>
#include <exception>
#include <string>
#include <memory>
#include <thread>
#include <utility>
#include <vector>

#include "common.hpp"

namespace nrgk {

struct BlockRef {
  long long off;  // element offset in its arena
  int rows, cols; // leading dimension == cols
};
// Block table of one operator: (row sector, column sector) -> block, iterated in key order like the
// reference's std::map-based qOperator (qOperator.hpp:18-158).  Entries live in one sorted vector:
// the producers add keys in ascending order (an append), lookups are binary searches, and copying a
// table (one per recorded shell) is a single allocation.
class OpMap {
 public:
  using key_type = std::pair<int, int>;
  using value_type = std::pair<key_type, BlockRef>;
  using iterator = std::vector<value_type>::iterator;
  using const_iterator = std::vector<value_type>::const_iterator;
  iterator begin() { return v_.begin(); }
  iterator end() { return v_.end(); }
  const_iterator begin() const { return v_.begin(); }
  const_iterator end() const { return v_.end(); }
  size_t size() const { return v_.size(); }
  bool empty() const { return v_.empty(); }
  void clear() { v_.clear(); }
  void reserve(size_t n) { v_.reserve(n); }
  iterator find(const key_type& k) {
    auto it = lower(k);
    return it != v_.end() && it->first == k ? it : v_.end();
  }
  const_iterator find(const key_type& k) const {
    auto it = std::lower_bound(v_.begin(), v_.end(), k, [](const value_type& a, const key_type& b) { return a.first < b; });
    return it != v_.end() && it->first == k ? it : v_.end();
  }
  size_t count(const key_type& k) const { return find(k) != end() ? 1 : 0; }
  const BlockRef& at(const key_type& k) const {
    auto it = find(k);
    if (it == end()) throw std::out_of_range("OpMap::at");
    return it->second;
  }
  BlockRef& operator[](const key_type& k) {
    if (v_.empty() || v_.back().first < k) {  // in-order insertion
      v_.push_back({k, BlockRef{0, 0, 0}});
      return v_.back().second;
    }
    auto it = lower(k);
    if (it == v_.end() || it->first != k) it = v_.insert(it, {k, BlockRef{0, 0, 0}});
    return it->second;
  }

 private:
  iterator lower(const key_type& k) {
    return std::lower_bound(v_.begin(), v_.end(), k, [](const value_type& a, const key_type& b) { return a.first < b; });
  }
  std::vector<value_type> v_;
};

// A site model (impurity or bath) in the reference's model concept.
struct SiteData {
  int S = 0, nqn = 0;
  std::vector<std::vector<int>> qn;
  std::vector<std::vector<double>> E;
  std::vector<double> chi;
  int nops = 0;
  std::vector<std::map<std::pair<int, int>, std::vector<double>>> ops;
  int dim(int s) const { return (int)E[s].size(); }
  static SiteData from_desc(const nrgk_site_desc* d);
};

struct Obs;

struct Chain {
  Ctx* ctx = nullptr;
  hipStream_t st = nullptr;
  SiteData imp, bath;
  size_t K = 1024;
  int mode = NRGK_MODE_MIXED;
  int cnt = 0, cnt_min = 0;
  size_t no_kept = 0;
  double errbar = 1e-8;
  // a state within tie_window of the cut (an eigenvalue error of that size could flip its keep or
  // discard decision) is counted in near_ties by every truncation
  double tie_window = 2e-10;
  double last_cut = 0.0;
  int near_ties = 0;
  bool keep_h = false;
  enum Phase { FRESH, ASSEMBLED, TRUNCATED } phase = FRESH;

  // previous (kept) basis, read by the next assembly
  std::vector<std::vector<int>> pre_qn;
  std::vector<int> pre_kept;
  std::vector<int> pre_eoff;
  const double* d_eprev = nullptr;
  DevBuf eprev_own;
  std::vector<OpMap> fprev;
  DevBuf fbuf[2];
  int fcur = 0;

  // bath on the device
  DevBuf d_bathE, d_bathF;
  std::vector<int> bath_eoff;
  std::vector<OpMap> bathF;

  // current iteration
  std::vector<std::vector<int>> cur_qn;
  std::vector<std::vector<std::pair<int, int>>> coupled;  // (ii, bb) per new sector
  std::vector<std::vector<int>> sub_kidx;                 // first row of each coupled sub-block
  std::vector<int> cur_n, cur_eoff;
  std::vector<long long> cur_off;
  long long tot_elems = 0;
  long long N_tot = 0;
  int nmax = 0;
  std::vector<int> kept;  // after update_internal_state

  DevBuf dH, dW, dX, dX2, dP, dS, dR, dE, dU, dY, dC, dQ, d_rrv;
  DevBuf dlam_tmp, dlam_raw, dlam_shift[2];
  int lcur = 0;
  DevBuf d_asm, d_eig, d_misc, d_rot, d_gemm;
  DevBuf d_flags, d_sigma_bits, d_sigma, d_stats, d_kept, d_trunc, d_conv, d_eig_act, d_cid, d_sidx, d_offmax, d_dgram, d_dgv;
  int oa_steps = 0, oa_unconverged = 0;
  bool fail_on_unconverged = false;  // nrgk_eigh_batched: an unconverged block is an error

  // host mirrors
  std::vector<double> all_eig;
  DevBuf dlam_prev;      // the previous iteration's raw eigenvalues, while all_eig is still unsorted
  long long N_prev = 0;
  std::vector<double> relgs;
  std::vector<double> h_last_shift;  // shifted eigenvalues of the previous basis (host copy, lazily)

  // diagnostics
  float ms[4] = {0, 0, 0, 0};
  // launches and host synchronisations of the last step (add_bath_site to the end of
  // update_internal_state, observable updates and records between them included)
  long long step_l0 = 0, step_s0 = 0, last_launches = 0, last_syncs = 0;
  int sweeps = 0, maxcluster = 0, unresolved = 0;
  hipEvent_t ev[6];
  std::vector<std::vector<double>> h_keep;  // assembled blocks (keep_h)

  // wide blocks: FP64 Householder tridiagonalisation on side stream 0, eigenvectors when needed
  // (wide.cpp)
  std::vector<HtBlock> ht;          // widest first
  std::vector<int> wide_sectors;    // sector of each ht entry
  DevBuf d_ht, d_wide_ws, d_wide_s, d_wide_q;
  DevBuf d_wide_A, d_wide_V;  // the working matrices (then V_p T_p) and Householder vectors, row stride ld
  double* wide_A = nullptr;
  double* wide_V = nullptr;
  enum WideState { WIDE_NONE, WIDE_EIGENVALUES, WIDE_KEPT, WIDE_ALL } wide_state = WIDE_NONE;
  bool wide_joined = true;
  // the host worker that enqueues the wide blocks' column loop on side stream 0 while this thread runs
  // the batched solve; its launches, timing runs and failure are taken over when it is joined
  std::thread wide_thread;
  std::exception_ptr wide_err;
  long long wide_launches = 0;
  // every eigenvector of every block is formed in add_bath_site (observables and FDM shells read
  // full U blocks); otherwise the wide blocks' kept columns only, after the truncation
  bool need_full_u = false;
  void ensure_side_streams();
  void wide_begin(const std::vector<int>& sectors);
  void wide_join();
  void wide_vectors(bool all);
  bool is_wide(int sector) const;

  // plain-mode handles (rocSOLVER / rocBLAS), created on first use
  void* rocblas = nullptr;
  // blocks wider than the batched solver's LDS panel: rocSOLVER on side streams (plain.cpp)
  static constexpr int N_SIDE = 4;
  hipStream_t st_big[N_SIDE] = {};
  void* rocblas_big[N_SIDE] = {};
  hipEvent_t ev_big_ready = nullptr;           // H assembled (main stream)
  hipEvent_t ev_big_done[N_SIDE] = {};         // a side stream's blocks done
  DevBuf d_big_work[N_SIDE], d_big_info;
  int n_big = 0;
  unsigned big_used = 0;  // side streams given blocks in this step (bit k: st_big[k])

  // per-kernel-family timing (HIP events on the chain's stream around runs of consecutive launches
  // of one family) with the algorithmic flops and bytes of those launches; off unless enabled
  struct KStat {
    double ms = 0, flops = 0, bytes = 0;
    long long launches = 0;
  };
  // small uploads and readbacks of the chain's stream go through page-locked staging; sync()
  // drains the stream and recycles it
  Staging stg;
  void sync();
  // host data to a persistent device buffer, without blocking the host (staged, then copied)
  template <class T>
  T* upload_to(DevBuf& buf, const T* src, size_t n, size_t byte_off = 0) {
    buf.ensure(byte_off + std::max<size_t>(1, n) * sizeof(T));
    T* dst = reinterpret_cast<T*>(buf.as<char>() + byte_off);
    if (n) {
      const T* d = stg.put(src, n, st);
      NRGK_HIP(hipMemcpyAsync(dst, d, n * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    return dst;
  }
  bool prof_on = false;
  std::map<std::string, KStat> kstats;
  // event pairs recorded around each run, resolved at the next sync() (recording never blocks)
  struct ProfRun {
    const char* family;
    hipEvent_t e0, e1;
    long long launches;
    double flops, bytes;
  };
  std::vector<hipEvent_t> ev_pool;
  std::vector<ProfRun> prof_pending;
  std::vector<ProfRun> wide_prof;  // the worker's timing runs, merged into prof_pending at the join
  void wide_thread_join();
  hipEvent_t prof_open = nullptr;
  hipEvent_t take_event();
  std::string prof_only;  // a single family to time (empty: all)
  void prof_begin(const char* family, hipStream_t s = nullptr);
  void prof_end(const char* family, long long launches, double flops, double bytes, hipStream_t s = nullptr);
  void prof_resolve();

  Chain(Ctx* c, const nrgk_site_desc* imp, const nrgk_site_desc* bath, size_t K, int mode);
  // the eigensolver alone, without sites (nrgk_eigh_batched)
  Chain(Ctx* c, int mode);
  // eigenpairs of the given row-major symmetric blocks (host arrays packed one after another):
  // ascending eigenvalues into w, eigenvectors as the columns of each block into U
  void solve_blocks(const std::vector<int>& dims, const double* H, double* w, double* U);
  // lays out blocks as an iteration's sectors and uploads H (mirrored from its upper triangle)
  void prepare_blocks(const std::vector<int>& dims, const double* H);
  int mixed_max_override = -1;  // the autotuner's crossover under test (-1: the context's)
  ~Chain();
  void set_bath(const nrgk_site_desc* b);
  void upload_bath();
  void import_impurity();
  void create_next_basis();
  void assemble(const std::vector<double>& t, double rescale);
  void eigensolve();
  void eigensolve_mixed();
  void eigensolve_plain();
  void add_bath_site(const double* t, int nt, double rescale);
  void update_internal_state();
  void rotate_fdag();
  void set_state(const nrgk_site_desc* st, int cnt);
  void ensure_all_eig();
  void run_gemms(std::vector<GemmProb>& probs);

 private:
  void init_device();
};

// plain FP64 mode (plain.cpp)
void plain_gemms(Chain& c, const std::vector<GemmProb>& probs);
void plain_destroy(Chain& c);
// eigenpairs of the given sectors by rocSOLVER dsyevd on the side stream, started once the main
// stream has reached this point; large_blocks_join makes the main stream wait for them
void large_blocks_begin(Chain& c, const std::vector<int>& sectors);
void large_blocks_join(Chain& c);
void large_blocks_check(Chain& c);  // after a synchronisation: the dsyevd infos
// FP64 dsyevd of the given sectors on the chain's stream (blocks the refinement left unconverged)
void solve_dense_fp64(Chain& c, const std::vector<int>& sectors);

}  // namespace nrgk
