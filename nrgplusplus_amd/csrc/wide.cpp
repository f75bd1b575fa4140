// The wide sector blocks of an iteration (wider than the batched Jacobi's crossover): their
// eigenvalues from a blocked FP64 Householder tridiagonalisation and bisection, on a side stream of
// the chain beside the batched solve of the other blocks; their eigenvectors by inverse iteration and
// WY back-transformation, only for the columns that are read (k_tridiag.hip).
//
// The reference diagonalises every sector with LAPACKE_dsyevd (qmatrix.hpp:375-404, called per
// sector at nrgcore.hpp:85-89), which always forms every eigenvector.  Here a chain without
// observables, FDM recording or an explicit request for U needs only the kept columns of the wide
// blocks (the f-dagger rotation reads U_i(:, 0:k_i), nrgcore.hpp:466-467), so their eigenvectors are
// formed after the truncation has fixed k_i; a later request for U (current_hamiltonQ) forms the rest.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "chain.hpp"
#include "kernels.hpp"

namespace nrgk {

static double env_knob(const char* name, double dflt) {
  const char* v = std::getenv(name);
  return v ? std::atof(v) : dflt;
}

void Chain::ensure_side_streams() {
  if (ev_big_ready) return;
  NRGK_HIP(hipEventCreateWithFlags(&ev_big_ready, hipEventDisableTiming));
  for (int k = 0; k < N_SIDE; k++) {
    NRGK_HIP(hipStreamCreateWithFlags(&st_big[k], hipStreamNonBlocking));
    NRGK_HIP(hipEventCreateWithFlags(&ev_big_done[k], hipEventDisableTiming));
  }
}

// Row stride of a wide block's working matrices: n rounded up to 16 doubles, so that every 64-column
// tile row is four whole 128-byte lines, and moved off multiples of 256 doubles (strides of a large
// power of two put a tile's rows on few HBM channels).
static int wide_ld(int n) {
  int ld = (n + 15) / 16 * 16;
  if (ld % 256 == 0) ld += 16;
  return ld;
}

// Tridiagonalisation and eigenvalues of the given sectors on side stream 0, started once the main
// stream has assembled H.  Writes the eigenvalues into dlam_raw; wide_join() makes the main stream
// wait for them.  Each block's working copy of H is in d_wide_A (its lower 64 x 64 tiles one after the
// other), its Householder vectors in d_wide_V (n x ld at the block's p_off, 128-byte aligned), the
// workspace in d_wide_ws.
void Chain::wide_begin(const std::vector<int>& sectors) {
  wide_thread_join();  // a worker of an earlier solve that was not joined (an error path)
  wide_err = nullptr;
  wide_sectors = sectors;
  std::stable_sort(wide_sectors.begin(), wide_sectors.end(), [&](int x, int y) { return cur_n[x] > cur_n[y]; });
  ht.clear();
  long long ws = 0, po = 0;
  int nmax = 1;
  for (int s : wide_sectors) {
    HtBlock b{};
    b.n = cur_n[s];
    b.T = (b.n + HT_TB - 1) / HT_TB;
    b.a_off = cur_off[s];
    b.ld = wide_ld(b.n);
    b.p_off = po;
    // the working matrix's lower tiles, or (after the reduction) V_p T_p at row stride ld
    po += (std::max((long long)b.n * b.ld, (long long)b.T * (b.T + 1) / 2 * 4096) + 31) / 32 * 32;
    b.w_off = ws;
    b.e_off = cur_eoff[s];
    ws += (ht_ws_size(b.n, b.T) + 31) / 32 * 32;
    nmax = std::max(nmax, b.n);
    ht.push_back(b);
  }
  wide_state = WIDE_NONE;
  if (ht.empty()) return;
  ensure_side_streams();
  hipStream_t s2 = st_big[0];
  d_wide_ws.ensure((size_t)ws * sizeof(double));
  d_wide_A.ensure((size_t)po * sizeof(double));
  d_wide_V.ensure((size_t)po * sizeof(double));
  d_ht.ensure(ht.size() * sizeof(HtBlock));
  // the descriptors go through the main stream's staging (the side stream starts after it)
  const HtBlock* dbl = stg.put(ht, st);
  NRGK_HIP(hipMemcpyAsync(d_ht.p, dbl, ht.size() * sizeof(HtBlock), hipMemcpyDeviceToDevice, st));
  NRGK_HIP(hipEventRecord(ev_big_ready, st));
  NRGK_HIP(hipStreamWaitEvent(s2, ev_big_ready, 0));
  const HtBlock* d = d_ht.as<HtBlock>();
  wide_A = d_wide_A.as<double>();
  wide_V = d_wide_V.as<double>();
  double* A = wide_A;
  double* V = wide_V;
  double* W = d_wide_ws.as<double>();
  launch_ht_tile_in(d, ht, dH.as<double>(), A, s2);
  for (const auto& b : ht) NRGK_HIP(hipMemsetAsync(V + b.p_off, 0, (size_t)b.n * b.ld * sizeof(double), s2));
  launch_ht_first(d, ht, A, W, s2);
  wide_state = WIDE_EIGENVALUES;
  wide_joined = false;
  wide_err = nullptr;
  wide_launches = 0;
  wide_prof.clear();
  // The column loop (two or three launches per column of the widest block, ~12 000 for a 4 300-row
  // block) is enqueued by a host worker, so that this thread reaches the batched solve of the other
  // blocks at once instead of after the enqueueing.  The worker touches only the side stream, its own
  // events and the block tables, which stay fixed until the join.
  wide_thread = std::thread([this, d, A, V, W, s2] {
    try {
      NRGK_HIP(hipSetDevice(ctx->device));
      const long long l0 = thread_counts().launches;
      auto want = [&](const char* f) { return prof_on && (prof_only.empty() || prof_only == f); };
      auto mark = [&](hipEvent_t& e) {
        NRGK_HIP(hipEventCreate(&e));
        NRGK_HIP(hipEventRecord(e, s2));
      };
      // algorithmic bytes of the symmetric products (the lower triangle of the trailing matrix read
      // once per column) and of the rank-2k updates (that triangle read and written once per panel).
      // With the profile restricted to "ht_symv", one column in HT_SYMV_SAMPLE has its k_ht_symv
      // launch timed by an event pair that the launch itself records at the kernel's start and end
      // (its duration and algorithmic bytes), so the kernel's own rate is measured without an event
      // between every pair of launches and without the enqueueing around it.
      constexpr int HT_SYMV_SAMPLE = 8;
      const bool symv_only = prof_on && prof_only == "ht_symv";
      const bool fam = want("tridiagonalise");
      hipEvent_t f0 = nullptr, f1 = nullptr;
      if (fam) mark(f0);
      double by = 0.0;
      long long launches = 0;
      for (int k = 0; k + 2 <= ht[0].n; k++) {
        double bsym = 0.0;
        if (prof_on)
          for (const auto& b : ht) {
            if (k > b.n - 2) break;
            const double m = b.n - k - 1;
            bsym += 8.0 * m * (m + 1) / 2;
            if (k % HT_NB == HT_NB - 1 || k == b.n - 2) by += 16.0 * m * (m + 1) / 2;
          }
        by += bsym;
        const bool sample = symv_only && k % HT_SYMV_SAMPLE == 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (sample) {
          NRGK_HIP(hipEventCreate(&e0));
          NRGK_HIP(hipEventCreate(&e1));
        }
        const int gs = launch_ht_symv(d, ht, k, A, V, W, s2, e0, e1);
        if (sample) {
          if (!gs) {  // nothing launched: the pair is recorded empty
            NRGK_HIP(hipEventRecord(e0, s2));
            NRGK_HIP(hipEventRecord(e1, s2));
          }
          wide_prof.push_back(ProfRun{"ht_symv", e0, e1, gs > 0, 0.0, bsym});
        }
        launches += gs > 0;
        launches += launch_ht_update(d, ht, k, A, V, W, s2) > 0;
        launches += launch_ht_syr2k(d, ht, k, A, V, W, s2) > 0;
      }
      if (fam) {
        mark(f1);
        wide_prof.push_back(ProfRun{"tridiagonalise", f0, f1, launches, 0.0, by});
      }
      hipEvent_t g0 = nullptr, g1 = nullptr;
      const bool fe = want("tridiag_eigenvalues");
      if (fe) mark(g0);
      launch_ht_stebz(d, ht, W, dlam_raw.as<double>(), s2);
      if (fe) {
        mark(g1);
        wide_prof.push_back(ProfRun{"tridiag_eigenvalues", g0, g1, 1, 0.0, 0.0});
      }
      NRGK_HIP(hipEventRecord(ev_big_done[0], s2));
      wide_launches = thread_counts().launches - l0;
    } catch (...) {
      wide_err = std::current_exception();
    }
  });
}

// The worker has enqueued everything (its launches and timing runs are taken over); its failure, if
// any, is rethrown here.
void Chain::wide_thread_join() {
  if (!wide_thread.joinable()) return;
  wide_thread.join();
  thread_counts().launches += wide_launches;
  wide_launches = 0;
  prof_pending.insert(prof_pending.end(), wide_prof.begin(), wide_prof.end());
  wide_prof.clear();
}

void Chain::wide_join() {
  wide_thread_join();
  if (wide_err) {
    std::exception_ptr e = wide_err;
    wide_err = nullptr;
    std::rethrow_exception(e);
  }
  if (ht.empty() || wide_joined) return;
  NRGK_HIP(hipStreamWaitEvent(st, ev_big_done[0], 0));
  wide_joined = true;
}

// Eigenvectors of the wide blocks on the main stream: all columns, or columns 0 .. kept - 1 (the
// truncation has run).  Inverse iteration on T into U's columns, then U = Q Z by the WY products.
void Chain::wide_vectors(bool all) {
  if (ht.empty() || wide_state == WIDE_NONE) return;
  if (wide_state == WIDE_ALL || (!all && wide_state == WIDE_KEPT)) return;
  wide_join();
  std::vector<HtBlock> bl;
  long long so = 0;
  for (size_t i = 0; i < ht.size(); i++) {
    HtBlock b = ht[i];
    const int s = wide_sectors[i];
    b.m = all ? b.n : std::min(b.n, kept.empty() ? b.n : kept[s]);
    if (b.m <= 0) continue;
    b.mpad = ht_stein_slots(b.m);
    b.s_off = so;
    so += 3LL * b.n * b.mpad;
    bl.push_back(b);
  }
  if (bl.empty()) {
    wide_state = all ? WIDE_ALL : WIDE_KEPT;
    return;
  }
  d_wide_s.ensure((size_t)so * sizeof(double));
  long long nq = 0;
  for (const auto& b : bl) nq += b.m;
  d_wide_q.ensure((size_t)(64 + 3 * nq) * sizeof(int));
  const HtBlock* d = stg.put(bl, st);
  double* W = d_wide_ws.as<double>();
  double* U = dU.as<double>();
  // clusters: eigenvalues within 5e-5 ||T|| of their neighbour.  Inverse iteration leaves vectors of
  // eigenvalues g apart orthogonal to about eps ||T|| / g, so that bound keeps the rest at ~4e-12
  // (1e-5 left the widest blocks of the two-channel K = 4096 flow at up to 6e-11, close to the 1e-10
  // bar; 5e-5 brings them to 1.3e-12 for 3.5 ms more orthogonalisation per step)
  const double ctol = env_knob("NRGK_WIDE_CLUSTER", 5e-5);
  prof_begin("tridiag_vectors");
  launch_ht_stein(d, bl, W, dlam_raw.as<double>(), U, d_wide_s.as<double>(), ctol, d_wide_q.as<int>() + 64,
                  d_wide_q.as<int>(), st);
  prof_end("tridiag_vectors", 6, 0.0, 0.0);  // the cluster list, two steps, three orthogonalisations
  // V_p T_p over each block's working matrix (no longer needed), then per panel, last first:
  // Y = V_p^T Z, Z -= (V_p T_p) Y, row-parallel (k_ht_wy_y, k_ht_wy_z)
  launch_ht_larft(d, bl, wide_V, W, wide_A, st);
  int pmax = 0;
  for (const auto& b : bl) pmax = std::max(pmax, (b.n - 1 + HT_NB - 1) / HT_NB);
  double* Ys = d_wide_s.as<double>();  // the inverse-iteration scratch is free again
  prof_begin("back_transform");
  long long nl = 0;
  for (int step = 0; step < pmax; step++) nl += launch_ht_wy(d, bl, step, wide_V, wide_A, U, Ys, st);
  prof_end("back_transform", nl, 0.0, 0.0);
  wide_state = all ? WIDE_ALL : WIDE_KEPT;
}

bool Chain::is_wide(int sector) const {
  return std::find(wide_sectors.begin(), wide_sectors.end(), sector) != wide_sectors.end() && !ht.empty();
}

}  // namespace nrgk
