// Plain FP64 GPU form (the comparator of the batched mixed path): every sector block is solved by
// its own rocSOLVER dsyevd call and every product is its own rocBLAS dgemm call, in FP64, without
// batching or graphs -- the reference's per-block LAPACKE_dsyevd / cblas_dgemm loop
// (nrgcore/include/nrgcore/nrgcore.hpp:85-89,466-467) moved to the GPU one call at a time.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>

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
  if (c.rocblas) rocblas_destroy_handle((rocblas_handle)c.rocblas);
  c.rocblas = nullptr;
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
    rocblas_status s = rocblas_dgemm(h, p.transB ? rocblas_operation_transpose : rocblas_operation_none,
                                     p.transA ? rocblas_operation_transpose : rocblas_operation_none, p.n, p.m, p.k,
                                     &alpha, p.B, p.ldb, p.A, p.lda, &beta, p.C, p.ldc);
    NRGK_REQUIRE(s == rocblas_status_success, NRGK_ERR_HIP, "rocblas_dgemm failed");
  }
  NRGK_HIP(hipStreamSynchronize(c.st));
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
    // symmetric input: the column-major view is the same matrix; eigenvectors come back as columns
    // of the column-major result, i.e. rows of the row-major buffer
    rocblas_status r = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, A, n,
                                        dlam_raw.as<double>() + cur_eoff[s], dlam_tmp.as<double>(), dinfo + s);
    NRGK_REQUIRE(r == rocblas_status_success, NRGK_ERR_HIP, "rocsolver_dsyevd failed");
    launch_transpose(A, dU.as<double>() + cur_off[s], n, st);
  }
  std::vector<int> info(cur_qn.size());
  d2h(info.data(), dinfo, info.size(), st);
  NRGK_HIP(hipStreamSynchronize(st));
  for (int v : info) NRGK_REQUIRE(v == 0, NRGK_ERR_NOCONV, "rocsolver_dsyevd: no convergence");
  sweeps = 0;
  maxcluster = 0;
  unresolved = 0;
}

}  // namespace nrgk
