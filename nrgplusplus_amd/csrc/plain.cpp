// Plain FP64 GPU form (the comparator of the batched mixed path): every sector block is solved by
// its own rocSOLVER dsyevd call and every product is its own rocBLAS dgemm call, in FP64, without
// batching or graphs -- the reference's per-block LAPACKE_dsyevd / cblas_dgemm loop
// (nrgcore/include/nrgcore/nrgcore.hpp:85-89,466-467) moved to the GPU one call at a time.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>

#include "chain.hpp"
#include "kernels.hpp"

namespace nrgk {

static rocblas_handle plain_handle(Chain& c) {
  if (!c.rocblas) {
    rocblas_handle h;
    NRGK_REQUIRE(rocblas_create_handle(&h) == rocblas_status_success, NRGK_ERR_HIP, "rocblas_create_handle failed");
    rocblas_set_stream(h, c.st);
    c.rocblas = h;
  }
  return (rocblas_handle)c.rocblas;
}

void plain_destroy(Chain& c) {
  if (c.st) (void)hipStreamSynchronize(c.st);
  if (c.rocblas) rocblas_destroy_handle((rocblas_handle)c.rocblas);
  c.rocblas = nullptr;
  for (int k = 0; k < Chain::N_SIDE; k++) {
    // the side stream drains before its handle goes
    if (c.st_big[k]) (void)hipStreamSynchronize(c.st_big[k]);
    if (c.rocblas_big[k]) rocblas_destroy_handle((rocblas_handle)c.rocblas_big[k]);
    c.rocblas_big[k] = nullptr;
    if (c.st_big[k]) (void)hipStreamDestroy(c.st_big[k]);
    c.st_big[k] = nullptr;
    if (c.ev_big_done[k]) (void)hipEventDestroy(c.ev_big_done[k]);
    c.ev_big_done[k] = nullptr;
  }
  if (c.ev_big_ready) (void)hipEventDestroy(c.ev_big_ready);
  c.ev_big_ready = nullptr;
}

// Row-major C = op(A) op(B) through column-major rocBLAS: C^T = op(B)^T op(A)^T.
void plain_gemms(Chain& c, const std::vector<GemmProb>& probs) {
  rocblas_handle h = plain_handle(c);
  for (const auto& p : probs) {
    if (p.m <= 0 || p.n <= 0) continue;
    if (p.beta != 0.0 && p.C0 != p.C)
      NRGK_HIP(hipMemcpy2DAsync(p.C, (size_t)p.ldc * sizeof(double), p.C0, (size_t)p.ldc0 * sizeof(double),
                                (size_t)p.n * sizeof(double), p.m, hipMemcpyDeviceToDevice, c.st));
    const double alpha = p.alpha, beta = p.beta;
    thread_counts().launches++;
    rocblas_status s = rocblas_dgemm(h, p.transB ? rocblas_operation_transpose : rocblas_operation_none,
                                     p.transA ? rocblas_operation_transpose : rocblas_operation_none, p.n, p.m, p.k,
                                     &alpha, p.B, p.ldb, p.A, p.lda, &beta, p.C, p.ldc);
    NRGK_REQUIRE(s == rocblas_status_success, NRGK_ERR_HIP, "rocblas_dgemm failed");
  }
  NRGK_HIP(stream_sync(c.st));
}

void Chain::eigensolve_plain() {
  rocblas_handle h = plain_handle(*this);
  const size_t bytes64 = std::max<long long>(1, tot_elems) * sizeof(double);
  dX.ensure(bytes64);
  dU.ensure(bytes64);
  dlam_raw.ensure(std::max<long long>(1, N_tot) * sizeof(double));
  dlam_tmp.ensure((std::max<long long>(1, N_tot) + 64) * sizeof(double));
  d_stats.ensure(std::max<size_t>(4, cur_qn.size()) * sizeof(int));
  int* dinfo = d_stats.as<int>();
  NRGK_HIP(hipMemsetAsync(dinfo, 0, cur_qn.size() * sizeof(int), st));
  for (size_t s = 0; s < cur_qn.size(); s++) {
    const int n = cur_n[s];
    if (n == 0) continue;
    double* A = dX.as<double>() + cur_off[s];
    NRGK_HIP(hipMemcpyAsync(A, dH.as<double>() + cur_off[s], (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, st));
    // the column-major view of the row-major block is its transpose: its lower triangle is the
    // row-major upper triangle that LAPACKE_dsyevd(ROW_MAJOR, 'U') reads (qmatrix.hpp:390-391).
    // Eigenvectors come back as columns of the column-major result, i.e. rows of the row-major buffer
    thread_counts().launches++;
    rocblas_status r = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, A, n,
                                        dlam_raw.as<double>() + cur_eoff[s], dlam_tmp.as<double>(), dinfo + s);
    NRGK_REQUIRE(r == rocblas_status_success, NRGK_ERR_HIP, "rocsolver_dsyevd failed");
    launch_transpose(A, dU.as<double>() + cur_off[s], n, st);
  }
  std::vector<int> info(cur_qn.size());
  d2h(info.data(), dinfo, info.size(), st);
  NRGK_HIP(stream_sync(st));
  for (int v : info) NRGK_REQUIRE(v == 0, NRGK_ERR_NOCONV, "rocsolver_dsyevd: no convergence");
  sweeps = 0;
  maxcluster = 0;
  unresolved = 0;
}


// The crossover of the batched solver: a block wider than its LDS-resident Jacobi panel (768 rows)
// is solved by rocSOLVER dsyevd -- Householder tridiagonalisation with blocked (GEMM) trailing
// updates, then divide and conquer, all in FP64 -- on side streams of the chain (largest blocks
// first, each to the stream with the least n^3 so far), while the main stream runs the batched FP32
// Jacobi and FP64 refinement of the other blocks.  The side streams write only the rows of their
// blocks in X (scratch), U and the raw eigenvalues, and have their own work and info arrays, so no
// two streams share memory.  All streams keep the default priority.
void large_blocks_begin(Chain& c, const std::vector<int>& sectors) {
  c.n_big = (int)sectors.size();
  c.big_used = 0;
  if (sectors.empty()) return;
  c.ensure_side_streams();
  for (int k = 0; k < Chain::N_SIDE; k++)
    if (!c.rocblas_big[k]) {
      rocblas_handle h;
      NRGK_REQUIRE(rocblas_create_handle(&h) == rocblas_status_success, NRGK_ERR_HIP, "rocblas_create_handle failed");
      rocblas_set_stream(h, c.st_big[k]);
      c.rocblas_big[k] = h;
      c.d_big_work[k].ctx = c.ctx;
    }
  c.d_big_info.ctx = c.ctx;
  std::vector<int> order(sectors);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return c.cur_n[x] > c.cur_n[y]; });
  double load[Chain::N_SIDE] = {0};
  c.d_big_info.ensure(sectors.size() * sizeof(int));
  NRGK_HIP(hipEventRecord(c.ev_big_ready, c.st));
  std::vector<std::vector<int>> mine(Chain::N_SIDE);
  for (int s : order) {
    const int k = (int)(std::min_element(load, load + Chain::N_SIDE) - load);
    load[k] += std::pow((double)c.cur_n[s], 3);
    mine[k].push_back(s);
  }
  int slot = 0;
  for (int k = 0; k < Chain::N_SIDE; k++) {
    if (mine[k].empty()) continue;
    c.big_used |= 1u << k;
    int nmax = 1;
    for (int s : mine[k]) nmax = std::max(nmax, c.cur_n[s]);
    c.d_big_work[k].ensure((size_t)nmax * sizeof(double));  // dsyevd's off-diagonal E
    NRGK_HIP(hipStreamWaitEvent(c.st_big[k], c.ev_big_ready, 0));
    rocblas_handle h = (rocblas_handle)c.rocblas_big[k];
    for (int s : mine[k]) {
      const int n = c.cur_n[s];
      double* A = c.dX.as<double>() + c.cur_off[s];
      NRGK_HIP(hipMemcpyAsync(A, c.dH.as<double>() + c.cur_off[s], (size_t)n * n * sizeof(double),
                              hipMemcpyDeviceToDevice, c.st_big[k]));
      // the row-major upper triangle (rocBLAS's column-major lower one) is read, as
      // LAPACKE_dsyevd(ROW_MAJOR, 'U') does; the eigenvectors come back in the rows of the
      // row-major buffer and are transposed into U's columns
      thread_counts().launches++;
    rocblas_status r = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, A, n,
                                          c.dlam_raw.as<double>() + c.cur_eoff[s], c.d_big_work[k].as<double>(),
                                          c.d_big_info.as<int>() + slot++);
      NRGK_REQUIRE(r == rocblas_status_success, NRGK_ERR_HIP, "rocsolver_dsyevd failed");
      launch_transpose(A, c.dU.as<double>() + c.cur_off[s], n, c.st_big[k]);
    }
    NRGK_HIP(hipEventRecord(c.ev_big_done[k], c.st_big[k]));
  }
}

// FP64 dsyevd of the given sectors on the chain's own stream (the blocks the batched refinement left
// unconverged): eigenvalues into dlam_raw, eigenvectors into U's columns, as the batched path
// writes them.  Returns after the stream has drained and the infos are checked.
void solve_dense_fp64(Chain& c, const std::vector<int>& sectors) {
  if (sectors.empty()) return;
  rocblas_handle h = plain_handle(c);
  int nmax = 1;
  for (int s : sectors) nmax = std::max(nmax, c.cur_n[s]);
  c.dlam_tmp.ensure(((size_t)std::max<long long>(1, c.N_tot) + nmax + 64) * sizeof(double));
  c.d_stats.ensure(std::max<size_t>(4, sectors.size()) * sizeof(int));
  int* dinfo = c.d_stats.as<int>();
  NRGK_HIP(hipMemsetAsync(dinfo, 0, sectors.size() * sizeof(int), c.st));
  for (size_t k = 0; k < sectors.size(); k++) {
    const int s = sectors[k], n = c.cur_n[s];
    double* A = c.dX.as<double>() + c.cur_off[s];
    NRGK_HIP(hipMemcpyAsync(A, c.dH.as<double>() + c.cur_off[s], (size_t)n * n * sizeof(double),
                            hipMemcpyDeviceToDevice, c.st));
    thread_counts().launches++;
    rocblas_status r = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, A, n,
                                        c.dlam_raw.as<double>() + c.cur_eoff[s], c.dlam_tmp.as<double>(), dinfo + k);
    NRGK_REQUIRE(r == rocblas_status_success, NRGK_ERR_HIP, "rocsolver_dsyevd failed");
    launch_transpose(A, c.dU.as<double>() + c.cur_off[s], n, c.st);
  }
  std::vector<int> info(sectors.size());
  d2h(info.data(), dinfo, info.size(), c.st);
  NRGK_HIP(stream_sync(c.st));
  for (int v : info) NRGK_REQUIRE(v == 0, NRGK_ERR_NOCONV, "rocsolver_dsyevd: no convergence");
}

void large_blocks_join(Chain& c) {
  if (!c.n_big) return;
  for (int k = 0; k < Chain::N_SIDE; k++)
    if (c.big_used >> k & 1u) NRGK_HIP(hipStreamWaitEvent(c.st, c.ev_big_done[k], 0));
}

void large_blocks_check(Chain& c) {
  if (!c.n_big) return;
  std::vector<int> info(c.n_big);
  d2h(info.data(), c.d_big_info.as<int>(), info.size(), c.st);
  NRGK_HIP(stream_sync(c.st));
  for (int v : info) NRGK_REQUIRE(v == 0, NRGK_ERR_NOCONV, "rocsolver_dsyevd: no convergence");
}

}  // namespace nrgk
