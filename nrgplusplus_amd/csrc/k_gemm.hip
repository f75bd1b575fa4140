// Batched variable-size FP64 GEMM: C = alpha * op(A) * op(B) + beta * C0 for a list of problems
// (row-major, per-problem transposes and leading dimensions).  It serves the refinement products
// (A X, X^T P, X^T X, X E), the operator rotations U_i^T T U_j (nrgcore.hpp:466-467,
// sysOperator.hpp:121-122) and the FDM products (fdmSpectrum.hpp:84-85,231-232).
//
// Every 64x64 output tile of every problem is one workgroup; tiles of all problems form a single
// grid (prefix sums in GemmProb::tile0), so one launch covers all uneven blocks of an iteration.
// A problem marked sym (a product known to be symmetric) gets only its upper tiles; the caller
// reads the upper triangle.
// The 64x16 / 16x64 operand slabs are staged in LDS (double-buffered) and each of the 256 threads
// keeps a 4x4 FP64 accumulator tile in registers.
#include "common.hpp"

namespace nrgk {

constexpr int GT = 64;   // tile edge
constexpr int GK = 16;   // k-slab

__device__ __forceinline__ int find_prob(const GemmProb* p, int n, int tile) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (p[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ void load_a(const GemmProb& P, int m0, int k0, double (*As)[GT + 1]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const int idx = tid + 256 * l;
    int mm, kk;
    if (!P.transA) { kk = idx % GK; mm = idx / GK; }
    else { mm = idx % GT; kk = idx / GT; }
    const int m = m0 + mm, k = k0 + kk;
    double v = 0.0;
    if (m < P.m && k < P.k) v = P.transA ? P.A[(long long)k * P.lda + m] : P.A[(long long)m * P.lda + k];
    As[kk][mm] = v;
  }
}
__device__ __forceinline__ void load_b(const GemmProb& P, int n0, int k0, double (*Bs)[GT + 1]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const int idx = tid + 256 * l;
    int nn, kk;
    if (!P.transB) { nn = idx % GT; kk = idx / GT; }
    else { kk = idx % GK; nn = idx / GK; }
    const int nc = n0 + nn, k = k0 + kk;
    double v = 0.0;
    if (nc < P.n && k < P.k) v = P.transB ? P.B[(long long)nc * P.ldb + k] : P.B[(long long)k * P.ldb + nc];
    Bs[kk][nn] = v;
  }
}

__global__ void __launch_bounds__(256) k_gemm64(const GemmProb* __restrict__ probs, int n_probs) {
  const int pi = find_prob(probs, n_probs, blockIdx.x);
  const GemmProb P = probs[pi];
  const int t = blockIdx.x - P.tile0;
  int tm, tn;
  if (P.sym) {  // upper tiles only (tm <= tn), row by row
    const int T = P.tiles_n;
    int rem = t;
    tm = 0;
    while (rem >= T - tm) {
      rem -= T - tm;
      tm++;
    }
    tn = tm + rem;
  } else {
    tm = t / P.tiles_n;
    tn = t % P.tiles_n;
  }
  const int m0 = tm * GT;
  const int n0 = tn * GT;
  __shared__ double As[2][GK][GT + 1];
  __shared__ double Bs[2][GK][GT + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
  const int nk = (P.k + GK - 1) / GK;
  if (nk > 0) {
    load_a(P, m0, 0, As[0]);
    load_b(P, n0, 0, Bs[0]);
  }
  __syncthreads();
  for (int ks = 0; ks < nk; ks++) {
    const int cur = ks & 1;
    if (ks + 1 < nk) {
      load_a(P, m0, (ks + 1) * GK, As[cur ^ 1]);
      load_b(P, n0, (ks + 1) * GK, Bs[cur ^ 1]);
    }
#pragma unroll
    for (int kk = 0; kk < GK; kk++) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[cur][kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[cur][kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int m = m0 + ty + 16 * i;
    if (m >= P.m) continue;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int nc = n0 + tx + 16 * j;
      if (nc >= P.n) continue;
      double v = P.alpha * acc[i][j];
      if (P.beta != 0.0) v += P.beta * P.C0[(long long)m * P.ldc0 + nc];
      P.C[(long long)m * P.ldc + nc] = v;
    }
  }
}

// Host helper: fills tile0/tiles_n, uploads the problem list and launches one grid.
int prepare_gemm(std::vector<GemmProb>& probs) {
  int tiles = 0;
  for (auto& p : probs) {
    p.tiles_n = (p.n + GT - 1) / GT;
    const int tm = (p.m + GT - 1) / GT;
    p.tile0 = tiles;
    NRGK_REQUIRE(!p.sym || p.m == p.n, NRGK_ERR_INVALID, "nrgk: symmetric product must be square");
    tiles += (p.m > 0 && p.n > 0) ? (p.sym ? tm * (tm + 1) / 2 : tm * p.tiles_n) : 0;
  }
  return tiles;
}

void launch_gemm(const GemmProb* d_probs, int n_probs, int tiles, hipStream_t s) {
  if (tiles <= 0 || n_probs <= 0) return;
  hipLaunchKernelGGL(k_gemm64, dim3(tiles), dim3(256), 0, s, d_probs, n_probs);
  NRGK_HIP(hipGetLastError());
}

}  // namespace nrgk
