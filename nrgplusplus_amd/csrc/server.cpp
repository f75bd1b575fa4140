// Request loop for impurity solves (no counterpart in the reference, whose drivers run one chain per
// process): spectral-function requests of different models, parameters, kept-state counts and chain
// lengths are admitted by their device-memory footprint and a slot count, each admitted request keeps
// its chain, observable and FDM shells resident on the device and runs on its own stream, so the
// iterations of all running requests interleave on the GPU; a request that finishes frees its slot
// and memory and the next one in line is admitted at once.  A failure inside one request (a status
// from any stage) ends that request only.
//
// Per request, the driver loop of examples/fdmSpectrumSiam/main.cpp:86-180: first site {V..V} with
// rescale 1, then n_sites Wilson sites (hopping(n, Lambda), rescale sqrt(Lambda) from the second),
// update_system_operator and a shell record whenever higher energies were discarded, the T = 0 FDM
// backward pass, and the log-Gaussian broadening of spec.py:27-44 on the requested frequencies.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "chain.hpp"
#include "kernels.hpp"
#include "observables.hpp"
#include "server.hpp"

namespace nrgk {

// Wilson-chain hopping t_n (examples/rgflowSIAM/main.cpp:5-10)
static double wilson_hopping(int n, double lambda) {
  const double x = n;
  return 0.5 * (1.0 + 1.0 / lambda) * (1.0 - std::pow(lambda, -x - 1.0)) /
         std::sqrt((1.0 - std::pow(lambda, -2.0 * x - 1.0)) * (1.0 - std::pow(lambda, -2.0 * x - 3.0)));
}

OwnedSite OwnedSite::copy(const nrgk_site_desc* d) {
  NRGK_REQUIRE(d && d->n_sectors > 0 && d->n_qn > 0 && d->qn && d->dims && d->chi, NRGK_ERR_INVALID,
               "nrgk_server_submit: incomplete site description");
  OwnedSite o;
  o.n_sectors = d->n_sectors;
  o.n_qn = d->n_qn;
  o.n_ops = d->n_ops;
  o.qn.assign(d->qn, d->qn + (size_t)d->n_sectors * d->n_qn);
  o.dims.assign(d->dims, d->dims + d->n_sectors);
  size_t ne = 0;
  for (int v : o.dims) ne += (size_t)v;
  o.energies.assign(d->energies, d->energies + ne);
  o.chi.assign(d->chi, d->chi + d->n_sectors);
  size_t nb = 0;
  if (d->n_ops > 0) {
    o.op_nblocks.assign(d->op_nblocks, d->op_nblocks + d->n_ops);
    for (int v : o.op_nblocks) nb += (size_t)v;
    o.op_bi.assign(d->op_bi, d->op_bi + nb);
    o.op_bj.assign(d->op_bj, d->op_bj + nb);
    size_t nd = 0;
    for (size_t b = 0; b < nb; b++) nd += (size_t)o.dims.at(o.op_bi[b]) * o.dims.at(o.op_bj[b]);
    o.op_data.assign(d->op_data, d->op_data + nd);
  }
  return o;
}

nrgk_site_desc OwnedSite::desc() const {
  nrgk_site_desc d{};
  d.n_sectors = n_sectors;
  d.n_qn = n_qn;
  d.qn = qn.data();
  d.dims = dims.data();
  d.energies = energies.data();
  d.chi = chi.data();
  d.n_ops = n_ops;
  d.op_nblocks = op_nblocks.data();
  d.op_bi = op_bi.data();
  d.op_bj = op_bj.data();
  d.op_data = op_data.data();
  return d;
}

// Device bytes a request is admitted against: the chain's per-iteration arenas (about a dozen n x n
// double arenas over the sector blocks) and one FDM shell per site (U blocks, eigenvalues, observable
// blocks).  The sector blocks of a chain at K kept states hold N = d_bath (K + 32) states and at most
// about sum_b n_b^2 = N^2 / 32 elements (the SIAM workloads have 0.02-0.03 N^2); the estimate admits
// against that bound, and the context's capacity remains the hard limit.
size_t Server::footprint(const Request& r) const {
  size_t dbath = 0;
  for (int v : r.bath.dims) dbath += (size_t)v;
  const double N = (double)dbath * ((double)r.max_kept + 32.0);
  const double sq = N * N / 32.0;
  const double arenas = 13.0 * 8.0 * sq;
  const double shells = (double)(r.n_sites + 1) * 8.0 * (3.0 * sq + N);
  return (size_t)(arenas + shells) + (size_t)(64u << 20);
}

Server::Server(Ctx* c, int slots, size_t budget) : ctx(c), max_slots(slots), hbm_budget(budget) {
  NRGK_REQUIRE(slots > 0 && slots <= 64, NRGK_ERR_INVALID, "nrgk_server_create: 1 to 64 slots");
  for (int i = 0; i < slots; i++) workers.emplace_back([this] { worker(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(m);
    stop = true;
  }
  cv.notify_all();
  for (auto& t : workers) t.join();
}

long long Server::submit(Request&& r) {
  NRGK_REQUIRE(r.max_kept > 0 && r.lambda > 1.0 && r.n_sites >= 0 && r.n_bins > 0 && r.n_omega >= 0 &&
                   r.broadening > 0 && std::isfinite(r.first_hopping),
               NRGK_ERR_INVALID, "nrgk_server_submit: bad request parameters");
  r.bytes = footprint(r);
  NRGK_REQUIRE(!hbm_budget || r.bytes <= hbm_budget, NRGK_ERR_NOMEM,
               "nrgk_server_submit: the request needs more device memory than the server's budget");
  std::lock_guard<std::mutex> lk(m);
  const long long id = next_id++;
  r.id = id;
  r.t_submit = std::chrono::steady_clock::now();
  auto p = std::make_shared<Request>(std::move(r));
  all[id] = p;
  pending.push_back(p);
  cv.notify_all();
  return id;
}

// the head of the line is admitted when a slot is free and its footprint fits the budget beside the
// running requests (first come, first served: a large request holds the line until memory frees)
bool Server::admissible_locked() const {
  if (pending.empty() || running >= max_slots) return false;
  return !hbm_budget || reserved + pending.front()->bytes <= hbm_budget;
}

void Server::worker() {
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(m);
      cv.wait(lk, [&] { return stop || admissible_locked(); });
      if (stop) return;
      r = pending.front();
      pending.pop_front();
      reserved += r->bytes;
      running++;
      peak_running = std::max(peak_running, running);
      peak_reserved = std::max(peak_reserved, reserved);
      r->t_start = std::chrono::steady_clock::now();
    }
    int status = NRGK_OK;
    std::string msg;
    try {
      solve(*r);
    } catch (const Error& e) {
      status = e.status;
      msg = e.what();
    } catch (const std::exception& e) {
      status = NRGK_ERR_INVALID;
      msg = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(m);
      r->status = status;
      r->message = msg;
      r->t_done = std::chrono::steady_clock::now();
      r->done = true;
      reserved -= r->bytes;
      running--;
      completed++;
    }
    cv.notify_all();
  }
}

void Server::solve(Request& r) {
  NRGK_HIP(hipSetDevice(ctx->device));
  const nrgk_site_desc di = r.impurity.desc(), db = r.bath.desc();
  Chain ch(ctx, &di, &db, r.max_kept, NRGK_MODE_MIXED);
  Obs obs(&ch, r.observable.n_ops, r.observable.op_nblocks.data(), r.observable.op_bi.data(), r.observable.op_bj.data(),
          r.observable.op_data.data());
  Fdm fdm(&ch, &obs, r.n_bins, 1e-10, 10.0);
  const int nop = r.impurity.n_ops;
  std::vector<double> t(nop, r.first_hopping);
  ch.add_bath_site(t.data(), nop, 1.0);
  obs.update();
  ch.update_internal_state();
  int min_it = 0;
  for (int k = 0; k < r.n_sites; k++) {
    const double tk = k == r.fail_at_site ? std::nan("") : wilson_hopping(k, r.lambda);
    std::fill(t.begin(), t.end(), tk);
    ch.add_bath_site(t.data(), nop, k == 0 ? 1.0 : std::sqrt(r.lambda));
    obs.update();
    if (ch.no_kept >= ch.K) fdm.record();  // checkHigherEnergyDiscarded (nrgcore.hpp:384-386)
    else min_it++;
    ch.update_internal_state();
    r.iterations++;
  }
  const int nops = obs.nops;
  std::vector<double> pos((size_t)nops * r.n_bins), neg((size_t)nops * r.n_bins);
  fdm.backward(r.lambda, -1, min_it + 1, pos.data(), neg.data(), nullptr, nullptr, 0, &r.shells);
  // summed over the operators (both spins), broadened on the device
  std::vector<double> x(r.n_bins), wp(r.n_bins, 0.0), wn(r.n_bins, 0.0);
  for (int i = 0; i < r.n_bins; i++) x[i] = 1e-10 * std::exp(i * 1. / fdm.delE);
  double tot = 0.0;
  for (int p = 0; p < nops; p++)
    for (int i = 0; i < r.n_bins; i++) {
      wp[i] += pos[(size_t)p * r.n_bins + i];
      wn[i] += neg[(size_t)p * r.n_bins + i];
      tot += pos[(size_t)p * r.n_bins + i] + neg[(size_t)p * r.n_bins + i];
    }
  r.total_weight = tot;
  r.A_pos.assign(r.n_omega, 0.0);
  r.A_neg.assign(r.n_omega, 0.0);
  if (r.n_omega > 0) {
    DevBuf d(ctx);
    const int nb = r.n_bins, no = r.n_omega;
    const size_t ns = broaden_scratch_doubles(nb, no);
    d.ensure(((size_t)3 * nb + 3 * no + ns + 8) * sizeof(double));
    double* dx = d.as<double>();
    double* dwp = dx + nb;
    double* dwn = dwp + nb;
    double* dom = dwn + nb;
    double* dap = dom + no;
    double* dan = dap + no;
    double* dscr = dan + no;
    h2d(dx, x.data(), nb, ch.st);
    h2d(dwp, wp.data(), nb, ch.st);
    h2d(dwn, wn.data(), nb, ch.st);
    h2d(dom, r.omega.data(), no, ch.st);
    launch_broaden(dx, dwp, nb, dom, no, r.broadening, dap, dscr, ch.st);
    launch_broaden(dx, dwn, nb, dom, no, r.broadening, dan, dscr, ch.st);
    d2h(r.A_pos.data(), dap, no, ch.st);
    d2h(r.A_neg.data(), dan, no, ch.st);
    NRGK_HIP(stream_sync(ch.st));
  }
}

}  // namespace nrgk
