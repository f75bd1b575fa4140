// Batched variable-size FP64 GEMM: C = alpha * op(A) * op(B) + beta * C0 for a list of problems
// (row-major, per-problem transposes and leading dimensions).  It serves the refinement products
// (A X, X^T P, X^T X, X E), the operator rotations U_i^T T U_j (nrgcore.hpp:466-467,
// sysOperator.hpp:121-122) and the FDM products (fdmSpectrum.hpp:84-85,231-232).
//
// Every 64x64 output tile of every problem is one workgroup; tiles of all problems form a single
// grid (prefix sums in GemmProb::tile0), so one launch covers all uneven blocks of an iteration.
// A problem marked sym (a product known to be symmetric) gets only its upper tiles; the caller
// reads the upper triangle.
// The 64x16 / 16x64 operand slabs are staged in LDS (double-buffered); the products run on the FP64
// matrix cores: each of the 4 waves owns a 32x32 quarter of the tile as 2x2 blocks of 16x16x4 steps.
#include "common.hpp"

namespace nrgk {

constexpr int GT = 64;   // tile edge
constexpr int GK = 16;   // k-slab
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int find_prob(const GemmProb* p, int n, int tile) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (p[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Operand slabs go global -> registers -> LDS: the registers of the next slab are fetched before
// the current slab's products, so the global latency overlaps them.  Thread t holds elements
// t + 256 l (l < GK/4) of the 64 x GK slab, walked along the contiguous direction of the operand.
constexpr int GL = GT * GK / 256;
__device__ __forceinline__ void fetch_a(const GemmProb& P, int m0, int k0, double* r) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int l = 0; l < GL; l++) {
    const int idx = tid + 256 * l;
    int mm, kk;
    if (!P.transA) { kk = idx % GK; mm = idx / GK; }
    else { mm = idx % GT; kk = idx / GT; }
    const int m = m0 + mm, k = k0 + kk;
    r[l] = (m < P.m && k < P.k) ? (P.transA ? P.A[(long long)k * P.lda + m] : P.A[(long long)m * P.lda + k]) : 0.0;
  }
}
__device__ __forceinline__ void store_a(const GemmProb& P, const double* r, double (*As)[GT + 1]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int l = 0; l < GL; l++) {
    const int idx = tid + 256 * l;
    if (!P.transA) As[idx % GK][idx / GK] = r[l];
    else As[idx / GT][idx % GT] = r[l];
  }
}
__device__ __forceinline__ void fetch_b(const GemmProb& P, int n0, int k0, double* r) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int l = 0; l < GL; l++) {
    const int idx = tid + 256 * l;
    int nn, kk;
    if (!P.transB) { nn = idx % GT; kk = idx / GT; }
    else { kk = idx % GK; nn = idx / GK; }
    const int nc = n0 + nn, k = k0 + kk;
    r[l] = (nc < P.n && k < P.k) ? (P.transB ? P.B[(long long)nc * P.ldb + k] : P.B[(long long)k * P.ldb + nc]) : 0.0;
  }
}
__device__ __forceinline__ void store_b(const GemmProb& P, const double* r, double (*Bs)[GT + 1]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int l = 0; l < GL; l++) {
    const int idx = tid + 256 * l;
    if (!P.transB) Bs[idx / GT][idx % GT] = r[l];
    else Bs[idx % GK][idx / GK] = r[l];
  }
}

__global__ void __launch_bounds__(256) k_gemm64(const GemmProb* __restrict__ probs, int n_probs) {
  // workgroup g runs on XCD g mod 8: XCD x takes the contiguous range of tiles [start(x), start(x+1)),
  // so the tiles of a row panel (and of one problem) share that XCD's L2
  const int T = gridDim.x, q = T >> 3, r = T & 7, x = blockIdx.x & 7;
  const int tile = x * q + min(x, r) + (blockIdx.x >> 3);
  const int pi = find_prob(probs, n_probs, tile);
  const GemmProb P = probs[pi];
  const int t = tile - P.tile0;
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
  // wave w: rows wm..wm+31, columns wn..wn+31 of the tile.  Lane l supplies A(m, k) for m = l%16 of
  // a 16-row block and k = l/16 of a 4-step, B(k, n) likewise; entry (i, j) of a 16x16 block lands
  // in lane 16 (i%4) + j, register i/4.
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  dbl4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
  const int nk = (P.k + GK - 1) / GK;
  double ra[GL], rb[GL];
  if (nk > 0) {
    fetch_a(P, m0, 0, ra);
    fetch_b(P, n0, 0, rb);
  }
  for (int ks = 0; ks < nk; ks++) {
    const int cur = ks & 1;
    // (the buffer written here was last read two slabs ago, before the previous barrier)
    store_a(P, ra, As[cur]);
    store_b(P, rb, Bs[cur]);
    __syncthreads();
    if (ks + 1 < nk) {
      fetch_a(P, m0, (ks + 1) * GK, ra);
      fetch_b(P, n0, (ks + 1) * GK, rb);
    }
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const double a0 = As[cur][kk + lk][wm + li], a1 = As[cur][kk + lk][wm + 16 + li];
      const double b0 = Bs[cur][kk + lk][wn + li], b1 = Bs[cur][kk + lk][wn + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int m = m0 + wm + 16 * bi + 4 * r + lk;
      if (m >= P.m) continue;
#pragma unroll
      for (int bj = 0; bj < 2; bj++) {
        const int nc = n0 + wn + 16 * bj + li;
        if (nc >= P.n) continue;
        double v = P.alpha * acc[bi][bj][r];
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
  NRGK_LAUNCH(k_gemm64, dim3(tiles), dim3(256), 0, s, d_probs, n_probs);
  NRGK_HIP(hipGetLastError());
}

}  // namespace nrgk
