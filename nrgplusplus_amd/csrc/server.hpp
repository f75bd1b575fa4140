// Request loop for impurity solves (server.cpp).
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

namespace nrgk {

// a site description with its own storage (requests outlive the caller's arrays)
struct OwnedSite {
  int n_sectors = 0, n_qn = 0, n_ops = 0;
  std::vector<int> qn, dims, op_nblocks, op_bi, op_bj;
  std::vector<double> energies, chi, op_data;
  static OwnedSite copy(const nrgk_site_desc* d);
  nrgk_site_desc desc() const;
};

struct Request {
  long long id = 0;
  OwnedSite impurity, bath, observable;
  size_t max_kept = 0;
  double lambda = 2.0, first_hopping = 0.0, broadening = 0.6;
  int n_sites = 0, n_bins = 0, n_omega = 0, fail_at_site = -1;
  std::vector<double> omega;
  size_t bytes = 0;  // admitted footprint
  // results
  bool done = false;
  int status = NRGK_OK;
  std::string message;
  int iterations = 0, shells = 0;
  double total_weight = 0.0;
  std::vector<double> A_pos, A_neg;
  std::chrono::steady_clock::time_point t_submit, t_start, t_done;
};

struct Server {
  Ctx* ctx;
  int max_slots;
  size_t hbm_budget;  // 0: the context's capacity alone
  std::mutex m;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Request>> pending;
  std::map<long long, std::shared_ptr<Request>> all;
  std::vector<std::thread> workers;
  bool stop = false;
  long long next_id = 1;
  size_t reserved = 0, peak_reserved = 0;
  int running = 0, peak_running = 0;
  long long completed = 0;

  Server(Ctx* c, int slots, size_t budget);
  ~Server();
  long long submit(Request&& r);
  size_t footprint(const Request& r) const;

 private:
  bool admissible_locked() const;
  void worker();
  void solve(Request& r);
};

}  // namespace nrgk
