// K9-K12: FDM spectral accumulation, reduced density matrix, thermodynamics and broadening.
//
// k_fdm_bin (fdmSpectrum::rhoDotOperators, nrgcore/include/dynamics/fdmSpectrum.hpp:207-277):
//   w+(v,v') = (B rho_j)(v,v') * B(v,v'),  w-(v,v') = (rho_i B)(v,v') * B(v,v'),
//   both zero where v and v' are kept states of this shell (the A = B^T kept x kept zeroing, :226-230);
//   bin = int(ln(1 + |scale (E_i[v] - E_j[v'])| / e_min) * delE) (:245-251); |w| > 1e-20 is added.
//   One workgroup per (operator block, row tile); the histogram is privatised in LDS and merged
//   into HBM with FP64 atomics once per workgroup.
// k_partial_trace (fdmSpectrum::setReduceDensityMatrix, :87-113).
// k_thermo (fdmThermodynamics::saveFinalData, nrgcore/include/static/fdmThermodynamics.hpp:154-226).
// k_broaden (examples/fdmSpectrumSiam/spec.py:27-44).
#include "common.hpp"

namespace nrgk {

constexpr int BIN_THREADS = 256;

__global__ void __launch_bounds__(BIN_THREADS) k_fdm_bin(const FdmBinItem* __restrict__ items, const double* __restrict__ B,
                                                         const double* __restrict__ RA, const double* __restrict__ AR,
                                                         const double* __restrict__ E, const double* __restrict__ bf,
                                                         double scale, double e_min, double delE, int nbins,
                                                         double* __restrict__ pos, double* __restrict__ neg,
                                                         double* __restrict__ spec_sum, double* __restrict__ signed_sum) {
  extern __shared__ double hist[];  // [2*nbins] when nbins is small enough, else unused
  const FdmBinItem it = items[blockIdx.x];
  const bool lds = nbins > 0 && (size_t)nbins * 2 * sizeof(double) <= 65536;
  if (lds)
    for (int b = threadIdx.x; b < 2 * nbins; b += blockDim.x) hist[b] = 0.0;
  __syncthreads();
  double* gp = pos + (size_t)it.op * nbins;
  double* gn = neg + (size_t)it.op * nbins;
  double ssum = 0.0, sgn = 0.0;  // |w| added to the bins; signed w+ + w- of every pair (sum rule)
  const long long tot = (long long)it.rows * it.nj;
  for (long long e = threadIdx.x; e < tot; e += blockDim.x) {
    const int v = it.row0 + (int)(e / it.nj), vp = (int)(e % it.nj);
    const double b = B[it.b_off + (long long)v * it.nj + vp];
    if (b == 0.0) continue;
    double ra, ar;
    if (it.diag_mode) {
      ra = b * bf[it.ej + vp];
      ar = bf[it.ei + v] * b;
    } else {
      if (v < it.kci && vp < it.kcj) continue;  // kept x kept excluded
      ra = vp < it.kcj ? RA[it.ra_off + (long long)v * it.kcj + vp] : 0.0;
      ar = v < it.kci ? AR[it.ar_off + (long long)v * it.nj + vp] : 0.0;
    }
    const double aa = ra * b, bbv = ar * b;
    sgn += aa + bbv;
    const int k = (int)(log(1.0 + (fabs(scale * (E[it.ei + v] - E[it.ej + vp])) / e_min)) * delE);
    if (k < nbins && k >= 0) {
      if (fabs(aa) > 1e-20) {
        if (lds) atomicAdd(&hist[k], fabs(aa));
        else atomicAdd(&gp[k], fabs(aa));
        ssum += fabs(aa);
      }
      if (fabs(bbv) > 1e-20) {
        if (lds) atomicAdd(&hist[nbins + k], fabs(bbv));
        else atomicAdd(&gn[k], fabs(bbv));
        ssum += fabs(bbv);
      }
    }
  }
  __syncthreads();
  if (lds)
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
      if (hist[b] != 0.0) atomicAdd(&gp[b], hist[b]);
      if (hist[nbins + b] != 0.0) atomicAdd(&gn[b], hist[nbins + b]);
    }
  for (int o = 32; o > 0; o >>= 1) {
    ssum += __shfl_xor(ssum, o, 64);
    sgn += __shfl_xor(sgn, o, 64);
  }
  if ((threadIdx.x & 63) == 0 && ssum != 0.0) atomicAdd(spec_sum, ssum);
  if ((threadIdx.x & 63) == 0 && sgn != 0.0) atomicAdd(&signed_sum[it.op], sgn);
}

// reducedRho[ii](it,itp) += sum_il rho_i[kidx + it + kii*il, kidx + itp + kii*il], contributions taken
// in the reference's order (sector i ascending, coupled sub-block order).  One workgroup per ii.
__global__ void __launch_bounds__(256) k_partial_trace(const PtItem* __restrict__ items, const PtContrib* __restrict__ cons,
                                                       const double* __restrict__ rho, double* __restrict__ red) {
  const PtItem it = items[blockIdx.x];
  const int k = it.kii;
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
    const int a = e / k, b = e % k;
    double r = 0.0;
    for (int c = it.c_begin; c < it.c_end; c++) {
      const PtContrib pc = cons[c];
      double aa = 0.0;
      for (int il = 0; il < pc.dbb; il++)
        aa += rho[pc.rho_off + (long long)(pc.kidx + a + k * il) * pc.ni + (pc.kidx + b + k * il)];
      r += aa;
    }
    red[it.r_off + e] = r;
  }
}

// Thermodynamics: for each temperature, per shell m: Z'_m = sum_{discarded} exp_reg((gs_m - E)/T),
// and the Boltzmann-weighted first and second moments of (E - gs_m).  One workgroup per T.
__device__ __forceinline__ double regulate_exp(double x) {
  if (x > 100) return exp(100.0);
  if (x < -100) return exp(-100.0);
  return exp(x);
}
__global__ void __launch_bounds__(256) k_thermo(const double* __restrict__ E, const int* __restrict__ shell_off,
                                                const double* __restrict__ gs, int n_shells, const double* __restrict__ T,
                                                double* __restrict__ z, double* __restrict__ m1, double* __restrict__ m2) {
  const int t = blockIdx.x;
  const double kT = T[t];
  __shared__ double red[3][4];
  for (int m = 0; m < n_shells; m++) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int e = shell_off[m] + threadIdx.x; e < shell_off[m + 1]; e += blockDim.x) {
      const double w = regulate_exp((gs[m] - E[e]) / kT);
      const double d = E[e] - gs[m];
      s0 += w;
      s1 += d * w;
      s2 += d * d * w;
    }
    for (int o = 32; o > 0; o >>= 1) {
      s0 += __shfl_xor(s0, o, 64);
      s1 += __shfl_xor(s1, o, 64);
      s2 += __shfl_xor(s2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = s0;
      red[1][threadIdx.x >> 6] = s1;
      red[2][threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      z[(size_t)t * n_shells + m] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
      m1[(size_t)t * n_shells + m] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
      m2[(size_t)t * n_shells + m] = red[2][0] + red[2][1] + red[2][2] + red[2][3];
    }
    __syncthreads();
  }
}

// Thermal expectation values of system operators: for each (temperature, operator) and shell m,
// sum over the shell's discarded states of exp_reg((gs_m - E)/T) * O_ll, O_ll aligned with E
// (fdmThermodynamics.hpp:191-206 before the w_m / Z'_m weighting).  One workgroup per (T, op).
__global__ void __launch_bounds__(256) k_thermo_ops(const double* __restrict__ E, const double* __restrict__ Q,
                                                    long long q_stride, const int* __restrict__ shell_off,
                                                    const double* __restrict__ gs, int n_shells,
                                                    const double* __restrict__ T, double* __restrict__ out) {
  const int t = blockIdx.x, q = blockIdx.y, nq = gridDim.y;
  const double kT = T[t];
  const double* Qq = Q + q * q_stride;
  __shared__ double red[4];
  for (int m = 0; m < n_shells; m++) {
    double s = 0;
    for (int e = shell_off[m] + threadIdx.x; e < shell_off[m + 1]; e += blockDim.x)
      s += regulate_exp((gs[m] - E[e]) / kT) * Qq[e];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[((size_t)t * nq + q) * n_shells + m] = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
  }
}

// Diagonal of the (i, i) blocks of the system operators, one workgroup per (operator, sector):
// dst[j] = src[j * (n + 1)] for j < n, zeros where the operator has no (i, i) block.
__global__ void __launch_bounds__(256) k_diag_gather(const DiagItem* __restrict__ items, const double* __restrict__ O,
                                                     double* __restrict__ out) {
  const DiagItem it = items[blockIdx.x];
  for (int j = threadIdx.x; j < it.n; j += blockDim.x)
    out[it.dst + j] = it.src >= 0 ? O[it.src + (long long)j * (it.n + 1)] : 0.0;
}

// Log-Gaussian broadening (examples/fdmSpectrumSiam/spec.py:27-44): out(w) = (1/w) sum_k weight_k
// exp(-(ln(w / x_k) / b - b / 4)^2) / (b sqrt(pi)).  The peaks are split into chunks of BROADEN_CHUNK:
// workgroup (frequency block, chunk) writes the chunk's partial sums to scratch[chunk][frequency], and
// k_broaden_sum adds the chunks in order (the same result whatever the launch shape).
constexpr int BROADEN_CHUNK = 512;
__global__ void __launch_bounds__(256) k_broaden_part(const double* __restrict__ x, const double* __restrict__ w, int n,
                                                      const double* __restrict__ omega, int no, double b,
                                                      double* __restrict__ part) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  const int k0 = blockIdx.y * BROADEN_CHUNK, k1 = min(n, k0 + BROADEN_CHUNK);
  __shared__ double sx[BROADEN_CHUNK], sw[BROADEN_CHUNK];
  for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    sx[k - k0] = x[k];
    sw[k - k0] = w[k];
  }
  __syncthreads();
  if (o >= no) return;
  const double lom = log(omega[o]);
  double s = 0.0;
  for (int k = 0; k < k1 - k0; k++) {
    if (sw[k] == 0.0) continue;
    const double y = -b * 0.25 + (lom - log(sx[k])) / b;
    s += sw[k] * exp(-y * y);
  }
  part[(long long)blockIdx.y * no + o] = s;
}
__global__ void __launch_bounds__(256) k_broaden_sum(const double* __restrict__ part, int n_chunks,
                                                     const double* __restrict__ omega, int no, double b,
                                                     double* __restrict__ out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= no) return;
  double s = 0.0;
  for (int c = 0; c < n_chunks; c++) s += part[(long long)c * no + o];
  out[o] = s / omega[o] / (b * sqrt(M_PI));
}

// dense transpose of square blocks (plain FP64 mode: rocSOLVER returns column-major eigenvectors)
__global__ void __launch_bounds__(256) k_transpose(const double* __restrict__ in, double* __restrict__ out, int n) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int j = ty; j < 32; j += 8)
    if (by + j < n && bx + tx < n) tile[j][tx] = in[(long long)(by + j) * n + bx + tx];
  __syncthreads();
  for (int j = ty; j < 32; j += 8)
    if (bx + j < n && by + tx < n) out[(long long)(bx + j) * n + by + tx] = tile[tx][j];
}

void launch_fdm_bin(const void* items, int n_items, const double* B, const double* RA, const double* AR,
                    const double* E, const double* bf, double scale, double e_min, double delE, int nbins, double* pos,
                    double* neg, double* spec_sum, double* signed_sum, hipStream_t s) {
  if (n_items <= 0) return;
  const size_t lds = (size_t)nbins * 2 * sizeof(double) <= 65536 ? (size_t)nbins * 2 * sizeof(double) : 0;
  NRGK_LAUNCH(k_fdm_bin, dim3(n_items), dim3(BIN_THREADS), lds, s, (const FdmBinItem*)items, B, RA, AR, E, bf,
                     scale, e_min, delE, nbins, pos, neg, spec_sum, signed_sum);
  NRGK_HIP(hipGetLastError());
}
void launch_partial_trace(const void* items, int n_items, const void* cons, const double* rho, double* red,
                          hipStream_t s) {
  if (n_items <= 0) return;
  NRGK_LAUNCH(k_partial_trace, dim3(n_items), dim3(256), 0, s, (const PtItem*)items, (const PtContrib*)cons, rho,
                     red);
  NRGK_HIP(hipGetLastError());
}
void launch_thermo(const double* E, const int* shell_off, const double* gs, int n_shells, const double* T, int nT,
                   double* z, double* m1, double* m2, hipStream_t s) {
  if (nT <= 0 || n_shells <= 0) return;
  NRGK_LAUNCH(k_thermo, dim3(nT), dim3(256), 0, s, E, shell_off, gs, n_shells, T, z, m1, m2);
  NRGK_HIP(hipGetLastError());
}
void launch_thermo_ops(const double* E, const double* Q, long long q_stride, int nq, const int* shell_off,
                       const double* gs, int n_shells, const double* T, int nT, double* out, hipStream_t s) {
  if (nT <= 0 || n_shells <= 0 || nq <= 0) return;
  NRGK_LAUNCH(k_thermo_ops, dim3(nT, nq), dim3(256), 0, s, E, Q, q_stride, shell_off, gs, n_shells, T, out);
  NRGK_HIP(hipGetLastError());
}
void launch_diag_gather(const DiagItem* items, int n_items, const double* O, double* out, hipStream_t s) {
  if (n_items <= 0) return;
  NRGK_LAUNCH(k_diag_gather, dim3(n_items), dim3(256), 0, s, items, O, out);
  NRGK_HIP(hipGetLastError());
}
size_t broaden_scratch_doubles(int n, int no) {
  return (size_t)((std::max(n, 1) + BROADEN_CHUNK - 1) / BROADEN_CHUNK) * (size_t)std::max(no, 1);
}
void launch_broaden(const double* x, const double* w, int n, const double* omega, int no, double b, double* out,
                    double* scratch, hipStream_t s) {
  if (no <= 0) return;
  const int nc = (std::max(n, 1) + BROADEN_CHUNK - 1) / BROADEN_CHUNK;
  NRGK_LAUNCH(k_broaden_part, dim3((no + 255) / 256, nc), dim3(256), 0, s, x, w, n, omega, no, b, scratch);
  NRGK_LAUNCH(k_broaden_sum, dim3((no + 255) / 256), dim3(256), 0, s, scratch, nc, omega, no, b, out);
  NRGK_HIP(hipGetLastError());
}
void launch_transpose(const double* in, double* out, int n, hipStream_t s) {
  if (n <= 0) return;
  NRGK_LAUNCH(k_transpose, dim3((n + 31) / 32, (n + 31) / 32), dim3(256), 0, s, in, out, n);
  NRGK_HIP(hipGetLastError());
}
size_t fdm_bin_item_size() { return sizeof(FdmBinItem); }

}  // namespace nrgk
