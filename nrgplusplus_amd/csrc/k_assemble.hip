// K1: per-(Q,Sz) sector Hamiltonian assembly (restates nrgcore::create_next_hamiltonians,
// nrgcore/include/nrgcore/nrgcore.hpp:146-230).
//
// One workgroup per global row; its threads stride over the row's columns, so every HBM write of
// the n x n FP64 block is coalesced.  Each element is computed by one thread from its own gather,
// with the reference's floating-point operations in the reference's order:
//   diagonal   h = E_prev[ii][it] * rescale + E_bath[bb][il]                    (nrgcore.hpp:181-183)
//   each term  h += ((F_prev(it,it') * F_bath(il',il)) * chi_bath[bb']) * t[p]  (nrgcore.hpp:204-215)
// applied in the order the reference's loop nest visits them, and with contraction off, so the
// block is bit-identical to the reference's for the same inputs.
#include "common.hpp"

namespace nrgk {

__device__ __forceinline__ int find_sub(const AsmSub* sub, int b, int e, int x) {
  int k = b;
  while (k + 1 < e && sub[k + 1].kidx <= x) k++;
  return k;
}

__global__ void __launch_bounds__(256) k_assemble(const AsmSector* __restrict__ sectors, int n_sectors,
                                                  const AsmSub* __restrict__ subs,
                                                  const AsmCon* __restrict__ cons,
                                                  const double* __restrict__ e_prev,
                                                  const double* __restrict__ e_bath,
                                                  const double* __restrict__ f_prev,
                                                  const double* __restrict__ f_bath, double rescale,
                                                  double* __restrict__ H) {
#pragma clang fp contract(off)
  const int grow = blockIdx.x;
  // sector of this row: binary search over row0
  int lo = 0, hi = n_sectors - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (sectors[mid].row0 <= grow) lo = mid;
    else hi = mid - 1;
  }
  const AsmSector sec = sectors[lo];
  const int r = grow - sec.row0;
  const int n = sec.n;
  if (r >= n) return;
  const int kr = find_sub(subs, sec.sub_begin, sec.sub_end, r) - sec.sub_begin;
  const AsmSub sr = subs[sec.sub_begin + kr];
  const int it_r = (r - sr.kidx) % sr.kii;
  const int il_r = (r - sr.kidx) / sr.kii;
  double* row = H + sec.h_off + (long long)r * n;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    const int kc = find_sub(subs, sec.sub_begin, sec.sub_end, c) - sec.sub_begin;
    const AsmSub sc = subs[sec.sub_begin + kc];
    const int it_c = (c - sc.kidx) % sc.kii;
    const int il_c = (c - sc.kidx) / sc.kii;
    double v = 0.0;
    if (r == c) v = e_prev[sr.e_off + it_r] * rescale + e_bath[sr.eb_off + il_r];
    for (int q = sec.con_begin; q < sec.con_end; q++) {
      const AsmCon e = cons[q];
      const bool direct = (e.k == kr && e.kp == kc);
      const bool mirror = (e.k == kc && e.kp == kr);
      if (!direct && !mirror) continue;
      double td = 0.0, tm = 0.0;
      if (direct) {
        double aa = f_prev[e.fp_off + (long long)it_r * e.fp_ld + it_c] * f_bath[e.fb_off + il_c * e.fb_ld + il_r];
        aa = aa * e.chi;
        td = aa * e.t;
      }
      if (mirror) {
        double aa = f_prev[e.fp_off + (long long)it_c * e.fp_ld + it_r] * f_bath[e.fb_off + il_r * e.fb_ld + il_c];
        aa = aa * e.chi;
        tm = aa * e.t;
      }
      if (direct && mirror) {
        // same (k,kp,p) loop: the direct write sits at loop position (it_r,il_r,it_c,il_c),
        // the mirror write at (it_c,il_c,it_r,il_r); on the diagonal the direct one is first.
        const bool direct_first = !(it_r > it_c || (it_r == it_c && il_r > il_c));
        if (direct_first) {
          v = v + td;
          v = v + tm;
        } else {
          v = v + tm;
          v = v + td;
        }
      } else if (direct) {
        v = v + td;
      } else {
        v = v + tm;
      }
    }
    row[c] = v;
  }
}

// FP32 working copy for the Jacobi sweeps: W = fp32(H) + sigma I with sigma from Gershgorin's bounds.
// Pass 1 per row: the absolute row sum and R_i - H_ii (R_i = off-diagonal absolute sum), reduced per
// block to rad = max row sum and low = max_i (R_i - H_ii) = -(Gershgorin lower bound).  The second
// is signed; it is kept in an order-preserving unsigned encoding.
__device__ __forceinline__ unsigned ordered_bits(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered_bits(unsigned o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// Gershgorin bounds per block: the largest absolute row sum (rad) and the largest R_i - H_ii
// (-(lower bound)), as order-preserving bits combined by atomicMax.  A workgroup takes GG_ROWS
// consecutive rows (a wave per row, GG_ROWS / 4 rows per wave) and issues one pair of atomics per
// block among them, so that the rows of a large block do not all contend for the same two words.
constexpr int GG_ROWS = 16;
__global__ void __launch_bounds__(256) k_gershgorin(const EigBlock* __restrict__ blocks, const int* __restrict__ row_block,
                                                    const int* __restrict__ row_local, int n_rows,
                                                    const double* __restrict__ H, unsigned int* __restrict__ sigma_bits) {
  __shared__ int sb[GG_ROWS];
  __shared__ unsigned srad[GG_ROWS], slow[GG_ROWS];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int q = 0; q < GG_ROWS / 4; q++) {
    const int lr = wv * (GG_ROWS / 4) + q;
    const int grow = blockIdx.x * GG_ROWS + lr;
    if (grow < n_rows) {  // uniform over the wave
      const int b = row_block[grow], r = row_local[grow];
      const EigBlock blk = blocks[b];
      const double* row = H + blk.off + (long long)r * blk.n;
      double s = 0.0;
      for (int c = lane; c < blk.n; c += 64) s += fabs(row[c]);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0) {
        const double d = row[r];
        // R_i - H_ii, rounded up so that the float bound is never below the exact one
        const double low = (s - fabs(d)) - d;
        float lf = (float)low;
        if ((double)lf < low) lf = nextafterf(lf, INFINITY);
        sb[lr] = b;
        srad[lr] = __float_as_uint((float)s);
        slow[lr] = ordered_bits(lf);
      }
    } else if (lane == 0) {
      sb[lr] = -1;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int cur = -1;
    unsigned mr = 0, ml = 0;
    for (int lr = 0; lr <= GG_ROWS; lr++) {
      const int b = lr < GG_ROWS ? sb[lr] : -1;
      if (b != cur && cur >= 0) {
        atomicMax(&sigma_bits[2 * cur], mr);
        atomicMax(&sigma_bits[2 * cur + 1], ml);
      }
      if (b < 0) break;
      if (b != cur) {
        cur = b;
        mr = srad[lr];
        ml = slow[lr];
      } else {
        mr = max(mr, srad[lr]);
        ml = max(ml, slow[lr]);
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_to_fp32_shift(const EigBlock* __restrict__ blocks, const int* __restrict__ row_block,
                                                       const int* __restrict__ row_local, const double* __restrict__ H,
                                                       const unsigned int* __restrict__ sigma_bits,
                                                       float* __restrict__ sigma_out, float* __restrict__ W) {
  const int grow = blockIdx.x;
  const int b = row_block[grow];
  const int r = row_local[grow];
  const EigBlock blk = blocks[b];
  // sigma = max(0, -lower bound) + 0.05 rad: W = H + sigma I is positive definite with its smallest
  // eigenvalue at least 0.05 rad, and as small a shift as the bound allows (a large shift compresses
  // the relative gaps of the singular values that one-sided Jacobi converges on)
  const float rad = __uint_as_float(sigma_bits[2 * b]);
  const float low = from_ordered_bits(sigma_bits[2 * b + 1]);
  const float sig = rad > 1e-20f ? fmaxf(low, 0.0f) + 0.05f * rad : 1.0f;
  const double* row = H + blk.off + (long long)r * blk.n;
  float* wrow = W + blk.off + (long long)r * blk.n;  // symmetric: row-major == column-major
  for (int c = threadIdx.x; c < blk.n; c += blockDim.x) {
    float v = (float)row[c];
    if (c == r) v += sig;
    wrow[c] = v;
  }
  if (r == 0 && threadIdx.x == 0) sigma_out[b] = sig;
}

void launch_assemble(const AsmSector* sectors, int n_sectors, int n_rows, const AsmSub* subs, const AsmCon* cons,
                     const double* e_prev, const double* e_bath, const double* f_prev, const double* f_bath,
                     double rescale, double* H, hipStream_t s) {
  if (n_rows <= 0) return;
  NRGK_LAUNCH(k_assemble, dim3(n_rows), dim3(256), 0, s, sectors, n_sectors, subs, cons, e_prev, e_bath,
                     f_prev, f_bath, rescale, H);
  NRGK_HIP(hipGetLastError());
}

void launch_shift_convert(const EigBlock* blocks, const int* row_block, const int* row_local, int n_rows,
                          const double* H, unsigned int* sigma_bits, float* sigma_out, float* W, hipStream_t s) {
  if (n_rows <= 0) return;
  NRGK_LAUNCH(k_gershgorin, dim3((n_rows + GG_ROWS - 1) / GG_ROWS), dim3(256), 0, s, blocks, row_block, row_local,
                     n_rows, H, sigma_bits);
  NRGK_HIP(hipGetLastError());
  NRGK_LAUNCH(k_to_fp32_shift, dim3(n_rows), dim3(256), 0, s, blocks, row_block, row_local, H, sigma_bits,
                     sigma_out, W);
  NRGK_HIP(hipGetLastError());
}

}  // namespace nrgk
