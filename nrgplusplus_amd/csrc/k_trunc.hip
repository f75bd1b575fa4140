// K7: truncation (restates nrgcore::discard_higher_energies, nrgcore/include/nrgcore/nrgcore.hpp:269-327).
//
// The reference sorts all N_tot eigenvalues and keeps, per sector, every state with
// E < sorted[K] + 1e-8 (strict, K 0-based), then shifts all energies by the minimum if and only if
// it truncated.  sorted[K] is found without a global sort: a radix select over the order-preserving
// 64-bit key of an FP64 value, eight passes of one 8-bit digit each (per-wave histograms in LDS, then
// the digit whose prefix range holds rank K).  The result is the exact element value, and the kept
// counts are lower_bound(sorted[K] + 1e-8) per block (each block's eigenvalues are ascending); both
// are bit-identical to the reference for identical eigenvalues.
#include "common.hpp"

namespace nrgk {


__device__ __forceinline__ unsigned long long okey(double x) {
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double okey_inv(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}
// number of elements < v (lower_bound) in an ascending list
__device__ __forceinline__ int count_lt(const double* a, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int m = (lo + hi) >> 1;
    if (a[m] < v) lo = m + 1;
    else hi = m;
  }
  return lo;
}

template <class T, class Op>
__device__ T block_reduce(T v, T* sh, Op op) {
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  T r = sh[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = op(r, sh[w]);
  __syncthreads();
  return r;
}

constexpr long long TRUNC_LDS = 16384;  // eigenvalues staged in LDS up to this count (128 KB)

__global__ void __launch_bounds__(1024) k_truncate(const int* __restrict__ bn, const int* __restrict__ beoff, int S,
                                                   const double* __restrict__ lam, long long N_tot, long long K,
                                                   double errbar, double tie_window, int* __restrict__ kept,
                                                   TruncOut* __restrict__ out, double* __restrict__ lam_shift) {
  __shared__ double shd[16];
  __shared__ long long shl[16];
  extern __shared__ double lds_lam[];  // all eigenvalues, when they fit (the searches then run in LDS)
  const int tid = threadIdx.x;
  const bool in_lds = N_tot <= TRUNC_LDS;
  if (in_lds) {
    for (long long i = tid; i < N_tot; i += blockDim.x) lds_lam[i] = lam[i];
    __syncthreads();
  }
  const double* L = in_lds ? lds_lam : lam;
  double mn = 1e308, mx = -1e308;
  for (int b = tid; b < S; b += blockDim.x)
    if (bn[b] > 0) {
      mn = fmin(mn, lam[beoff[b]]);
      mx = fmax(mx, lam[beoff[b] + bn[b] - 1]);
    }
  mn = block_reduce(mn, shd, [](double a, double c) { return fmin(a, c); });
  mx = block_reduce(mx, shd, [](double a, double c) { return fmax(a, c); });
  const bool trunc = N_tot > K;
  double e_cut = mx;
  if (trunc) {
    // radix select of the element of rank K (0-based) among all N_tot keys
    __shared__ int hist[16][256];  // one histogram per wave
    __shared__ int s_digit;
    __shared__ long long s_rank;
    const int wv = tid >> 6, lane = tid & 63, nwv = blockDim.x >> 6;
    unsigned long long prefix = 0, mask = 0;
    long long rank = K;
    for (int shift = 56; shift >= 0; shift -= 8) {
      for (int e = tid; e < 16 * 256; e += blockDim.x) (&hist[0][0])[e] = 0;
      __syncthreads();
      for (long long i = tid; i < N_tot; i += blockDim.x) {
        const unsigned long long k = okey(L[i]);
        if ((k & mask) == prefix) atomicAdd(&hist[wv][(k >> shift) & 255], 1);
      }
      __syncthreads();
      if (tid < 256) {
        int c = 0;
        for (int w = 0; w < nwv; w++) c += hist[w][tid];
        // inclusive scan of the 256 bin counts: within each wave, then across the four waves
        int incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(incl, o, 64);
          if (lane >= o) incl += y;
        }
        hist[0][tid] = incl;  // bins of this wave, read back below after the barrier
        hist[1][tid] = c;
      }
      __syncthreads();
      if (tid < 256) {
        long long before = hist[0][tid] - hist[1][tid];
        for (int w = 0; w < (tid >> 6); w++) before += hist[0][64 * w + 63];
        if (before <= rank && rank < before + hist[1][tid]) {
          s_digit = tid;
          s_rank = rank - before;
        }
      }
      __syncthreads();
      prefix |= (unsigned long long)s_digit << shift;
      mask |= 255ull << shift;
      rank = s_rank;
      __syncthreads();
    }
    e_cut = okey_inv(prefix);
  }
  const double x = e_cut + errbar;
  long long nk = 0;
  for (int b = tid; b < S; b += blockDim.x) {
    const int k = trunc ? count_lt(L + beoff[b], bn[b], x) : bn[b];
    kept[b] = k;
    nk += k;
  }
  nk = block_reduce(nk, shl, [](long long a, long long d) { return a + d; });
  // near-ties: states whose keep/discard decision an eigenvalue error of tie_window could flip
  long long nt = 0, nf = 0;
  for (long long i = tid; i < N_tot; i += blockDim.x) {
    if (trunc) nt += fabs(L[i] - x) <= tie_window ? 1 : 0;
    nf += isfinite(L[i]) ? 0 : 1;
  }
  nt = block_reduce(nt, shl, [](long long a, long long d) { return a + d; });
  nf = block_reduce(nf, shl, [](long long a, long long d) { return a + d; });
  // shift every eigenvalue (not only the kept ones) when truncating (nrgcore.hpp:320-325); the
  // per-block eigenvalue ranges are contiguous prefix offsets covering [0, N_tot)
  for (long long i = tid; i < N_tot; i += blockDim.x) lam_shift[i] = trunc ? lam[i] - mn : lam[i];
  if (tid == 0) {
    out->e_cut = e_cut;
    out->e_min = mn;
    out->e_max = mx;
    out->truncated = trunc ? 1 : 0;
    out->n_kept = (int)nk;
    out->near_ties = (int)nt;
    out->nonfinite = (int)nf;
  }
}

void set_trunc_lds_limit() {
  NRGK_HIP(hipFuncSetAttribute((const void*)k_truncate, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(TRUNC_LDS * sizeof(double))));
}

void launch_truncate(const int* bn, const int* beoff, int S, const double* lam, long long N_tot, long long K,
                     double errbar, double tie_window, int* kept, TruncOut* out, double* lam_shift, hipStream_t s) {
  const size_t lds = N_tot <= TRUNC_LDS ? (size_t)std::max<long long>(1, N_tot) * sizeof(double) : 0;
  NRGK_LAUNCH(k_truncate, dim3(1), dim3(1024), lds, s, bn, beoff, S, lam, N_tot, K, errbar, tie_window, kept, out,
              lam_shift);
  NRGK_HIP(hipGetLastError());
}

}  // namespace nrgk
