// K3-K5: the eigensolver of the wide sector blocks (above the batched Jacobi's crossover), replacing
// per-block LAPACKE_dsyevd (nrgcore/include/utils/qmatrix.hpp:375-404) for those blocks.
//
// 1. Blocked Householder tridiagonalisation in FP64 (the lower-triangle form: reflector H_k = I -
//    tau_k v_k v_k^T annihilates column k below its subdiagonal).  Columns are reduced in panels of
//    HT_NB; inside a panel the trailing matrix is not touched: each column's symmetric product
//    y = A v is formed from the panel-start matrix and corrected by the panel's V and W, and the
//    panel ends with one rank-2k update A -= V W^T + W V^T.  All wide blocks of an iteration advance
//    column by column in the same launches:
//      k_ht_symv    one workgroup per 64 x 64 tile of the trailing lower triangle: the tile's row
//                   sums (A v) and column sums (A^T v over the strict lower part), v^T A v and the
//                   panel's V^T v, W^T v shares -- the HBM-bound bulk (each element read once per
//                   column, the lower triangle only)
//      k_ht_update  one workgroup per 64-row tile-row: reduces those partials to w, stores v and w,
//                   and forms the next column with the panel corrections and its norm
//      k_ht_syr2k   the panel's rank-2k update of the trailing lower triangle; the tiles of the
//                   next column also extract it
//    Every scalar a later launch needs (the reflector of column k from the norm partials) is
//    recomputed by each workgroup from the same partials in the same order, so all agree bit for bit
//    and no launch waits on a single-workgroup reduction.
// 2. Eigenvalues of T by bisection on Sturm counts (one thread per eigenvalue, T in LDS), ascending
//    as dsyevd returns them.
// 3. Eigenvectors of T by inverse iteration with partial-pivoting LU of T - lambda I (one thread per
//    cluster of eigenvalues closer than ctol, the cluster's members orthogonalised against each other
//    by modified Gram-Schmidt), for the columns that are needed: all of them, or only the kept ones.
// 4. Back-transformation U = Q Z in WY form, Q_p = I - V_p T_p V_p^T per panel (k_ht_larft forms
//    V_p T_p), as two batched FP64 GEMMs per panel on the matrix cores (k_gemm64).
#include <cmath>

#include "common.hpp"
#include "kernels.hpp"

namespace nrgk {

__device__ __forceinline__ double ht_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// tiles of the lower triangle of the trailing matrix that starts at row/column s, and its tile-rows
__host__ __device__ __forceinline__ int ht_rows(int T, int s) {
  const int m = T - s / HT_TB;
  return m > 0 ? m : 0;
}
__host__ __device__ __forceinline__ int ht_tiles(int T, int s) {
  const int m = ht_rows(T, s);
  return m * (m + 1) / 2;
}
// the block's panel ends at column k: the rank-2k update follows
__host__ __device__ __forceinline__ bool ht_panel_end(int n, int k) { return k % HT_NB == HT_NB - 1 || k == n - 2; }
// absolute index of lower tile (I, J), J <= I
__device__ __forceinline__ long long ht_tidx(int I, int J) { return (long long)I * (I + 1) / 2 + J; }
// lower tile t (0-based, row by row) of an m x m tile triangle
__device__ __forceinline__ void ht_tile_of(int t, int& Ip, int& Jp) {
  int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= t) i++;
  while (i * (i + 1) / 2 > t) i--;
  Ip = i;
  Jp = t - i * (i + 1) / 2;
}

// Workgroup g of a column-k launch over the blocks [0, nblk) (sorted by n, widest first; a block takes
// part while k <= n - 2): its block b and task t.  KIND 0: tiles, 1: tile-rows, 2: tiles of the blocks
// whose panel ends at k.
template <int KIND>
__device__ __forceinline__ bool ht_find(const HtBlock* bl, int nblk, int k, int g, int& b, int& t) {
  for (int i = 0; i < nblk; i++) {
    const int n = bl[i].n;
    if (k > n - 2) break;
    if (KIND == 2 && !ht_panel_end(n, k)) continue;
    const int cnt = KIND == 1 ? ht_rows(bl[i].T, k + 1) : ht_tiles(bl[i].T, k + 1);
    if (g < cnt) {
      b = i;
      t = g;
      return true;
    }
    g -= cnt;
  }
  return false;
}
// the same counts on the host (grid sizes)
int ht_grid(const std::vector<HtBlock>& bl, int k, int kind) {
  int g = 0;
  for (const auto& b : bl) {
    if (k > b.n - 2) break;
    if (kind == 2 && !ht_panel_end(b.n, k)) continue;
    g += kind == 1 ? ht_rows(b.T, k + 1) : ht_tiles(b.T, k + 1);
  }
  return g;
}

// The reflector of column k (dlarfg's form): from alpha = A(k+1, k) and the squared norm of the
// entries below it (the tile-rows' partials, summed in tile-row order).
struct HtRefl {
  double alpha, beta, tau, sc;
};
__device__ HtRefl ht_reflector(const HtWs& w, int n, int T, int k) {
  const int s = k + 1;
  double xn2 = 0.0;
  const double* ss = w.ss + (long long)(k & 1) * T;
  for (int I = s / HT_TB; I < T; I++) xn2 += ss[I];
  const double* u = w.u + (long long)(k & 1) * n;
  HtRefl r;
  r.alpha = u[s];
  if (xn2 == 0.0) {
    r.beta = r.alpha;
    r.tau = 0.0;
    r.sc = 0.0;
  } else {
    const double nrm = sqrt(r.alpha * r.alpha + xn2);
    r.beta = r.alpha >= 0.0 ? -nrm : nrm;
    r.tau = (r.beta - r.alpha) / r.beta;
    r.sc = 1.0 / (r.alpha - r.beta);
  }
  return r;
}
// entry r of v_k (v[k+1] = 1, zero above)
__device__ __forceinline__ double ht_v(const double* u, const HtRefl& R, int s, int n, int r) {
  return (r < s || r >= n) ? 0.0 : (r == s ? 1.0 : u[r] * R.sc);
}

// Column 0 of every block: u = A(1:n, 0), d_0 = A(0, 0), squared-norm partials of A(2:n, 0).
__global__ void __launch_bounds__(64) k_ht_first(const HtBlock* __restrict__ bl, int nblk, const double* __restrict__ Aa,
                                                 double* __restrict__ Wa) {
  int g = blockIdx.x, b = -1;
  for (int i = 0; i < nblk; i++) {
    if (g < bl[i].T) {
      b = i;
      break;
    }
    g -= bl[i].T;
  }
  if (b < 0) return;
  const HtBlock B = bl[b];
  const int n = B.n;
  const HtWs w = ht_ws(const_cast<double*>(Wa) + B.w_off, n, B.T);
  const double* A = Aa + B.a_off;
  const int r = g * HT_TB + threadIdx.x;
  double sq = 0.0;
  if (r < n) {
    const double a = A[(long long)r * n];
    if (r == 0) w.d[0] = a;
    else w.u[r] = a;
    if (r >= 2) sq = a * a;
  }
  sq = ht_wave_sum(sq);
  if (threadIdx.x == 0) w.ss[g] = sq;  // column 0: buffer 0
}

__global__ void __launch_bounds__(256) k_ht_symv(const HtBlock* __restrict__ bl, int nblk, int k,
                                                 const double* __restrict__ Aa, const double* __restrict__ Va,
                                                 double* __restrict__ Wa) {
  int b, t;
  if (!ht_find<0>(bl, nblk, k, blockIdx.x, b, t)) return;
  const HtBlock B = bl[b];
  const int n = B.n, T = B.T, s = k + 1, t0 = s / HT_TB;
  const HtWs w = ht_ws(Wa + B.w_off, n, T);
  const double* A = Aa + B.a_off;
  const double* V = Va + B.a_off;
  const double* u = w.u + (long long)(k & 1) * n;
  int Ip, Jp;
  ht_tile_of(t, Ip, Jp);
  const int I = t0 + Ip, J = t0 + Jp;
  __shared__ HtRefl R;
  __shared__ double vr[HT_TB], vc[HT_TB];
  __shared__ double red[4][16][HT_TB + 1];
  __shared__ double cred[4][HT_TB];
  __shared__ double sp2[2];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  if (tid == 0) R = ht_reflector(w, n, T, k);
  __syncthreads();
  if (tid < 64) vr[tid] = ht_v(u, R, s, n, I * HT_TB + tid);
  else if (tid < 128) vc[tid - 64] = ht_v(u, R, s, n, J * HT_TB + tid - 64);
  __syncthreads();
  // wave wv: rows 16 wv .. 16 wv + 15 of the tile, lane: column (one 512-byte row segment per load)
  const int c = J * HT_TB + lane;
  const bool diag = I == J;
  double a[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = I * HT_TB + wv * 16 + i;
    const bool ok = r < n && c < n && r >= s && c >= s && (!diag || c <= r);
    a[i] = ok ? A[(long long)r * n + c] : 0.0;
  }
  double colacc = 0.0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = I * HT_TB + wv * 16 + i;
    red[wv][i][lane] = a[i] * vc[lane];
    colacc += (!diag || c < r) ? a[i] * vr[wv * 16 + i] : 0.0;
  }
  cred[wv][lane] = colacc;
  __syncthreads();
  const long long ix = ht_tidx(I, J);
  double myv = 0.0;
  if (tid < 64) {
    double rs = 0.0;
    for (int l = 0; l < HT_TB; l++) rs += red[tid >> 4][tid & 15][l];
    w.rp[ix * 64 + tid] = rs;
    myv = rs * vr[tid];
  } else if (tid < 128) {
    const int q = tid - 64;
    const double cs = (cred[0][q] + cred[1][q]) + (cred[2][q] + cred[3][q]);
    w.cp[ix * 64 + q] = cs;
    myv = cs * vc[q];
  }
  myv = ht_wave_sum(myv);
  if (lane == 0 && wv < 2) sp2[wv] = myv;
  __syncthreads();
  if (tid == 0) w.sp[ix] = sp2[0] + sp2[1];
  if (diag) {
    // this tile-row's shares of V^T v and W^T v over the panel's earlier columns
    const int j0 = (k / HT_NB) * HT_NB, cc = k - j0;
    if (tid < cc) {
      double acc = 0.0;
      for (int l = 0; l < HT_TB; l++) {
        const int r = I * HT_TB + l;
        if (r >= s && r < n) acc += V[(long long)r * n + j0 + tid] * vr[l];
      }
      w.qv[I * HT_NB + tid] = acc;
    } else if (tid >= 64 && tid - 64 < cc) {
      const int j = tid - 64;
      double acc = 0.0;
      for (int l = 0; l < HT_TB; l++) {
        const int r = I * HT_TB + l;
        if (r >= s && r < n) acc += w.W[(long long)r * HT_NB + j] * vr[l];
      }
      w.qw[I * HT_NB + j] = acc;
    }
  }
}

// w(r) = tau (y(r) - V(r,:) W^T v - W(r,:) V^T v) - (tau/2) (w^T v) v(r), y from the tile partials
__device__ __forceinline__ double ht_wrow(const HtWs& w, const double* V, int n, int T, int k, int j0, int cc,
                                          const HtRefl& R, const double* u, const double* Vv, const double* Wv,
                                          double wvs, int r) {
  const int s = k + 1, t0 = s / HT_TB, I = r / HT_TB, q = r - I * HT_TB;
  double y = 0.0;
  for (int J = t0; J <= I; J++) y += w.rp[ht_tidx(I, J) * 64 + q];
  for (int Ir = I; Ir < T; Ir++) y += w.cp[ht_tidx(Ir, I) * 64 + q];
  double corr = 0.0;
  for (int j = 0; j < cc; j++) corr += V[(long long)r * n + j0 + j] * Wv[j] + w.W[(long long)r * HT_NB + j] * Vv[j];
  return R.tau * (y - corr) - 0.5 * R.tau * wvs * ht_v(u, R, s, n, r);
}

__global__ void __launch_bounds__(256) k_ht_update(const HtBlock* __restrict__ bl, int nblk, int k,
                                                   const double* __restrict__ Aa, double* __restrict__ Va,
                                                   double* __restrict__ Wa) {
  int b, t;
  if (!ht_find<1>(bl, nblk, k, blockIdx.x, b, t)) return;
  const HtBlock B = bl[b];
  const int n = B.n, T = B.T, s = k + 1, t0 = s / HT_TB, I = t0 + t;
  const int j0 = (k / HT_NB) * HT_NB, cc = k - j0;
  const bool pend = ht_panel_end(n, k);
  const HtWs w = ht_ws(Wa + B.w_off, n, T);
  const double* A = Aa + B.a_off;
  double* V = Va + B.a_off;
  const double* u = w.u + (long long)(k & 1) * n;
  __shared__ HtRefl R;
  __shared__ double Vv[HT_NB], Wv[HT_NB];
  __shared__ double sred[256];
  __shared__ double sh_wv, sh_w1;
  const int tid = threadIdx.x;
  if (tid == 0) R = ht_reflector(w, n, T, k);
  // v^T A v over the active tiles: a fixed partition and tree, the same in every workgroup
  const int m = T - t0, NTa = m * (m + 1) / 2;
  double acc = 0.0;
  for (int q = tid; q < NTa; q += 256) {
    int Ip, Jp;
    ht_tile_of(q, Ip, Jp);
    acc += w.sp[ht_tidx(t0 + Ip, t0 + Jp)];
  }
  sred[tid] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) sred[tid] += sred[tid + o];
    __syncthreads();
  }
  if (tid < cc) {
    double a = 0.0, bq = 0.0;
    for (int Ir = t0; Ir < T; Ir++) {
      a += w.qv[Ir * HT_NB + tid];
      bq += w.qw[Ir * HT_NB + tid];
    }
    Vv[tid] = a;
    Wv[tid] = bq;
  }
  __syncthreads();
  if (tid == 0) {
    double vw = 0.0;
    for (int j = 0; j < cc; j++) vw += Vv[j] * Wv[j];
    sh_wv = R.tau * (sred[0] - 2.0 * vw);
  }
  __syncthreads();
  const double wvs = sh_wv;
  if (tid == 0 && !pend) sh_w1 = ht_wrow(w, V, n, T, k, j0, cc, R, u, Vv, Wv, wvs, s);
  const int r = I * HT_TB + tid;
  double wr = 0.0, vrr = 0.0;
  const bool own = tid < 64 && r < n && r >= s;
  if (own) {
    wr = ht_wrow(w, V, n, T, k, j0, cc, R, u, Vv, Wv, wvs, r);
    vrr = ht_v(u, R, s, n, r);
    w.W[(long long)r * HT_NB + cc] = wr;
    V[(long long)r * n + k] = vrr;
  }
  if (t == 0 && tid == 0) {
    w.tau[k] = R.tau;
    w.e[k] = R.beta;
  }
  __syncthreads();
  if (pend) return;
  // the next column (k + 1 = s) with the panel's corrections, its diagonal entry and norm partial
  const double w1 = sh_w1;
  double sq = 0.0;
  if (own) {
    double un = A[(long long)r * n + s];
    for (int j = 0; j < cc; j++)
      un -= V[(long long)r * n + j0 + j] * w.W[(long long)s * HT_NB + j] +
            w.W[(long long)r * HT_NB + j] * V[(long long)s * n + j0 + j];
    un -= vrr * w1 + wr;  // column k: V(r, k) W(s, k) + W(r, k) V(s, k), V(s, k) = 1
    if (r == s) w.d[s] = un;
    else w.u[(long long)(s & 1) * n + r] = un;
    if (r >= s + 2) sq = un * un;
  }
  if (tid < 64) {
    sq = ht_wave_sum(sq);
    if (tid == 0) w.ss[(long long)(s & 1) * T + I] = sq;
  }
}

// Rank-2k update of the trailing lower triangle (rows and columns >= s = k + 1) at the end of a panel:
// A -= V W^T + W V^T over the panel's columns; the tiles of tile-column s / HT_TB then extract column s.
__global__ void __launch_bounds__(256) k_ht_syr2k(const HtBlock* __restrict__ bl, int nblk, int k,
                                                  double* __restrict__ Aa, const double* __restrict__ Va,
                                                  double* __restrict__ Wa) {
  int b, t;
  if (!ht_find<2>(bl, nblk, k, blockIdx.x, b, t)) return;
  const HtBlock B = bl[b];
  const int n = B.n, T = B.T, s = k + 1, t0 = s / HT_TB;
  const int j0 = (k / HT_NB) * HT_NB, nbu = k - j0 + 1;
  const HtWs w = ht_ws(Wa + B.w_off, n, T);
  double* A = Aa + B.a_off;
  const double* V = Va + B.a_off;
  int Ip, Jp;
  ht_tile_of(t, Ip, Jp);
  const int I = t0 + Ip, J = t0 + Jp;
  // thread: rows 4 tr .. 4 tr + 3, columns 4 tc .. 4 tc + 3 of the tile
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  __shared__ double VI[16][HT_TB], WI[16][HT_TB], VJ[16][HT_TB], WJ[16][HT_TB];
  __shared__ double sqp[16];
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
  for (int jb = 0; jb < nbu; jb += 16) {
    // stage 16 panel columns of V and W for the tile's rows (I) and columns (J)
    for (int e = tid; e < 16 * HT_TB; e += 256) {
      const int jj = e >> 6, l = e & 63, j = jb + jj;
      const int ri = I * HT_TB + l, rj = J * HT_TB + l;
      const bool okj = j < nbu;
      VI[jj][l] = okj && ri < n ? V[(long long)ri * n + j0 + j] : 0.0;
      WI[jj][l] = okj && ri < n ? w.W[(long long)ri * HT_NB + j] : 0.0;
      VJ[jj][l] = okj && rj < n ? V[(long long)rj * n + j0 + j] : 0.0;
      WJ[jj][l] = okj && rj < n ? w.W[(long long)rj * HT_NB + j] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int jj = 0; jj < 16; jj++) {
      double vi[4], wi[4], vj[4], wj[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        vi[q] = VI[jj][4 * tr + q];
        wi[q] = WI[jj][4 * tr + q];
        vj[q] = VJ[jj][4 * tc + q];
        wj[q] = WJ[jj][4 * tc + q];
      }
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] += vi[i] * wj[j] + wi[i] * vj[j];
    }
    __syncthreads();
  }
  const bool diag = I == J;
  double sq = 0.0;
  const int cs = s - J * HT_TB;  // column s inside this tile (when J == t0)
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = I * HT_TB + 4 * tr + i;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int c = J * HT_TB + 4 * tc + j;
      if (r < n && c < n && r >= s && c >= s && (!diag || c <= r)) {
        const long long o = (long long)r * n + c;
        const double na = A[o] - acc[i][j];
        A[o] = na;
        if (c == s) {
          if (r == s) w.d[s] = na;
          else w.u[(long long)(s & 1) * n + r] = na;
          if (r >= s + 2) sq += na * na;
        }
      }
    }
  }
  if (J == t0) {
    // the threads holding column s: tc == cs / 4 (one per 4 rows); the tile-row's norm partial
    if (tc == (cs >> 2)) sqp[tr] = sq;
    __syncthreads();
    if (tid == 0) {
      double a = 0.0;
      for (int q = 0; q < 16; q++) a += sqp[q];
      w.ss[(long long)(s & 1) * T + I] = a;
    }
  }
}

// Eigenvalues of T (d, e) by bisection on the Sturm count, the j-th smallest by thread j; T is read
// from LDS (every thread walks the same entries).
__global__ void __launch_bounds__(256) k_ht_stebz(const HtBlock* __restrict__ bl, int nblk, const double* __restrict__ Wa,
                                                  double* __restrict__ lam) {
  extern __shared__ double lds[];
  int g = blockIdx.x, b = -1;
  for (int i = 0; i < nblk; i++) {
    const int cw = (bl[i].n + 255) / 256;
    if (g < cw) {
      b = i;
      break;
    }
    g -= cw;
  }
  if (b < 0) return;
  const HtBlock B = bl[b];
  const int n = B.n;
  const HtWs w = ht_ws(const_cast<double*>(Wa) + B.w_off, n, B.T);
  double* d = lds;
  double* e2 = lds + n;
  __shared__ double rlo[256], rhi[256], rnm[256];
  const int tid = threadIdx.x;
  double lo = 1e300, hi = -1e300, nm = 0.0;
  for (int i = tid; i < n; i += 256) {
    const double di = w.d[i];
    const double ei = i < n - 1 ? w.e[i] : 0.0, ep = i > 0 ? w.e[i - 1] : 0.0;
    d[i] = di;
    e2[i] = ei * ei;
    const double rad = fabs(ei) + fabs(ep);
    lo = fmin(lo, di - rad);
    hi = fmax(hi, di + rad);
    nm = fmax(nm, fabs(di) + rad);
  }
  rlo[tid] = lo;
  rhi[tid] = hi;
  rnm[tid] = nm;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) {
      rlo[tid] = fmin(rlo[tid], rlo[tid + o]);
      rhi[tid] = fmax(rhi[tid], rhi[tid + o]);
      rnm[tid] = fmax(rnm[tid], rnm[tid + o]);
    }
    __syncthreads();
  }
  const double tnorm = rnm[0];
  const double eps = 2.220446049250313e-16;
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, tnorm * tnorm);
  const double atol = 0.25 * eps * tnorm;
  const int j = g * 256 + tid;
  if (j >= n) return;
  double a = rlo[0] - 2.0 * eps * tnorm - pivmin, c = rhi[0] + 2.0 * eps * tnorm + pivmin;
  for (int it = 0; it < 200; it++) {
    const double mid = 0.5 * (a + c);
    if (c - a <= 2.0 * eps * fmax(fabs(a), fabs(c)) + atol || mid <= a || mid >= c) break;
    // Sturm count: eigenvalues below mid
    int cnt = 0;
    double q = d[0] - mid;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
    for (int i = 1; i < n; i++) {
      q = (d[i] - mid) - e2[i - 1] / q;
      if (fabs(q) < pivmin) q = -pivmin;
      cnt += q < 0.0;
    }
    if (cnt <= j) a = mid;
    else c = mid;
  }
  lam[B.e_off + j] = 0.5 * (a + c);
}

// deterministic start vector entries in [-1, 1)
__device__ __forceinline__ double ht_start(unsigned p, unsigned i) {
  unsigned long long x = ((unsigned long long)p << 32) ^ (unsigned long long)(i + 0x9E3779B9u);
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// One inverse-iteration step for eigenvalue lp: the iterate in column p of U (stride n) is replaced by
// (T - lp I)^-1 times it.  T - lp I = P L U with partial pivoting; the factorisation and the
// elimination of the iterate share one pass (the L factor is not kept), the upper factor's three
// diagonals go to the scratch f0, f1, f2 at entry i * mp + slot, then one backward pass.
__device__ void ht_inverse_step(const double* d, const double* e, int n, double lp, double piv0, double* U, int p,
                                double* f0, double* f1, double* f2, long long mp, int slot) {
  double cd = d[0] - lp, cu = n > 1 ? e[0] : 0.0;
  double xi = U[p];
  for (int i = 0; i < n - 1; i++) {
    const double sub = e[i], nd = d[i + 1] - lp, nu = i + 1 < n - 1 ? e[i + 1] : 0.0;
    const double xn = U[(long long)(i + 1) * n + p];
    const long long o = (long long)i * mp + slot;
    if (fabs(cd) >= fabs(sub)) {
      if (cd == 0.0) cd = piv0;
      const double ml = sub / cd;
      f0[o] = cd;
      f1[o] = cu;
      f2[o] = 0.0;
      U[(long long)i * n + p] = xi;
      xi = xn - ml * xi;
      cd = nd - ml * cu;
      cu = nu;
    } else {  // rows i and i + 1 exchanged
      const double ml = cd / sub;
      f0[o] = sub;
      f1[o] = nd;
      f2[o] = nu;
      U[(long long)i * n + p] = xn;
      xi = xi - ml * xn;
      cd = cu - ml * nd;
      cu = -ml * nu;
    }
  }
  if (cd == 0.0) cd = piv0;
  f0[(long long)(n - 1) * mp + slot] = cd;
  U[(long long)(n - 1) * n + p] = xi;
  double x1 = 0.0, x2 = 0.0;
  for (int i = n - 1; i >= 0; i--) {
    const long long o = (long long)i * mp + slot;
    double v = U[(long long)i * n + p];
    if (i + 1 < n) v -= f1[o] * x1;
    if (i + 2 < n) v -= f2[o] * x2;
    v /= f0[o];
    U[(long long)i * n + p] = v;
    x2 = x1;
    x1 = v;
  }
}

constexpr int HT_SMALL_CLUSTER = 8;  // clusters up to this size: one thread each
constexpr int HT_ITERS = 3;          // inverse-iteration steps per eigenvector

// the infinity norm of T (d, e in memory p)
__device__ __forceinline__ double ht_tnorm_entry(const double* d, const double* e, int n, int i) {
  return fabs(d[i]) + (i < n - 1 ? fabs(e[i]) : 0.0) + (i > 0 ? fabs(e[i - 1]) : 0.0);
}

// Eigenvectors of T for the eigenvalues 0 .. m-1 of each block (and the rest of a cluster that
// starts below m), into the columns of U.  Thread j starts a cluster when lambda_j is more than ctol
// above lambda_{j-1}.  A cluster of at most HT_SMALL_CLUSTER members is done by its thread: three
// inverse-iteration steps per member in turn, each followed by orthogonalisation against the members
// before it (modified Gram-Schmidt) and normalisation.  A larger cluster is queued for
// k_ht_stein_wide.  The LU factors live in interleaved scratch (entry i of slot j at i * mpad + j).
__global__ void __launch_bounds__(256) k_ht_stein(const HtBlock* __restrict__ bl, int nblk, const double* __restrict__ Wa,
                                                  const double* __restrict__ lam, double* __restrict__ Ua,
                                                  double* __restrict__ Sa, double ctol_rel, int* __restrict__ queue) {
  extern __shared__ double lds[];
  int g = blockIdx.x, b = -1;
  for (int i = 0; i < nblk; i++) {
    const int cw = bl[i].mpad / 256;
    if (g < cw) {
      b = i;
      break;
    }
    g -= cw;
  }
  if (b < 0) return;
  const HtBlock B = bl[b];
  const int n = B.n;
  const long long mp = B.mpad;
  const HtWs w = ht_ws(const_cast<double*>(Wa) + B.w_off, n, B.T);
  double* d = lds;
  double* e = lds + n;
  __shared__ double rnm[256];
  const int tid = threadIdx.x;
  double nm = 0.0;
  for (int i = tid; i < n; i += 256) {
    d[i] = w.d[i];
    e[i] = i < n - 1 ? w.e[i] : 0.0;
    nm = fmax(nm, ht_tnorm_entry(w.d, w.e, n, i));
  }
  rnm[tid] = nm;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) rnm[tid] = fmax(rnm[tid], rnm[tid + o]);
    __syncthreads();
  }
  const double tnorm = rnm[0];
  const int j = g * 256 + tid;
  if (j >= B.m) return;
  const double* L = lam + B.e_off;
  const double ctol = ctol_rel * tnorm;
  if (j > 0 && L[j] - L[j - 1] <= ctol) return;  // inside a cluster started below
  int j1 = j + 1;
  while (j1 < n && L[j1] - L[j1 - 1] <= ctol) j1++;
  if (j1 - j > HT_SMALL_CLUSTER) {
    const int q = atomicAdd(queue, 1);
    queue[1 + 2 * q] = b;
    queue[2 + 2 * q] = j;
    return;
  }
  double* U = Ua + B.a_off;
  double* f0 = Sa + B.s_off;
  double* f1 = f0 + (long long)n * mp;
  double* f2 = f1 + (long long)n * mp;
  // a pivot that vanishes is replaced by eps ||T|| (by 1 for T = 0, where any basis is one)
  const double piv0 = tnorm > 0.0 ? 2.220446049250313e-16 * tnorm : 1.0;
  for (int p = j; p < j1; p++) {
    const double lp = L[p];
    for (int i = 0; i < n; i++) U[(long long)i * n + p] = ht_start((unsigned)p, (unsigned)i);
    for (int it = 0; it < HT_ITERS; it++) {
      ht_inverse_step(d, e, n, lp, piv0, U, p, f0, f1, f2, mp, j);
      for (int q = j; q < p; q++) {
        double dt = 0.0;
        for (int i = 0; i < n; i++) dt += U[(long long)i * n + q] * U[(long long)i * n + p];
        for (int i = 0; i < n; i++) U[(long long)i * n + p] -= dt * U[(long long)i * n + q];
      }
      double nr = 0.0;
      for (int i = 0; i < n; i++) {
        const double v = U[(long long)i * n + p];
        nr += v * v;
      }
      const double inv = nr > 0.0 ? 1.0 / sqrt(nr) : 0.0;
      for (int i = 0; i < n; i++) U[(long long)i * n + p] *= inv;
    }
  }
}

// The queued clusters (more than HT_SMALL_CLUSTER members): one 64-thread workgroup per cluster, taken
// from the queue in strides of the grid.  Thread 0 runs the (sequential) inverse-iteration steps; the
// orthogonalisation and normalisation are spread over the workgroup, with a barrier between the
// phases.  Dot products are summed in a fixed order (per-thread strided sums, then a tree).
__global__ void __launch_bounds__(64) k_ht_stein_wide(const HtBlock* __restrict__ bl, const double* __restrict__ Wa,
                                                      const double* __restrict__ lam, double* __restrict__ Ua,
                                                      double* __restrict__ Sa, double ctol_rel,
                                                      const int* __restrict__ queue) {
  __shared__ double red[64];
  __shared__ double sh_tn;
  const int nq = queue[0];
  const int tid = threadIdx.x;
  for (int q = blockIdx.x; q < nq; q += gridDim.x) {
    const int b = queue[1 + 2 * q], j = queue[2 + 2 * q];
    const HtBlock B = bl[b];
    const int n = B.n;
    const long long mp = B.mpad;
    const HtWs w = ht_ws(const_cast<double*>(Wa) + B.w_off, n, B.T);
    double nm = 0.0;
    for (int i = tid; i < n; i += 64) nm = fmax(nm, ht_tnorm_entry(w.d, w.e, n, i));
    red[tid] = nm;
    __syncthreads();
    for (int o = 32; o > 0; o >>= 1) {
      if (tid < o) red[tid] = fmax(red[tid], red[tid + o]);
      __syncthreads();
    }
    if (tid == 0) sh_tn = red[0];
    __syncthreads();
    const double tnorm = sh_tn;
    const double* L = lam + B.e_off;
    const double ctol = ctol_rel * tnorm;
    int j1 = j + 1;
    while (j1 < n && L[j1] - L[j1 - 1] <= ctol) j1++;
    double* U = Ua + B.a_off;
    double* f0 = Sa + B.s_off;
    double* f1 = f0 + (long long)n * mp;
    double* f2 = f1 + (long long)n * mp;
    // a pivot that vanishes is replaced by eps ||T|| (by 1 for T = 0, where any basis is one)
  const double piv0 = tnorm > 0.0 ? 2.220446049250313e-16 * tnorm : 1.0;
    for (int p = j; p < j1; p++) {
      for (int i = tid; i < n; i += 64) U[(long long)i * n + p] = ht_start((unsigned)p, (unsigned)i);
      __syncthreads();
      for (int it = 0; it < HT_ITERS; it++) {
        if (tid == 0) ht_inverse_step(w.d, w.e, n, L[p], piv0, U, p, f0, f1, f2, mp, j);
        __syncthreads();
        for (int q2 = j; q2 < p; q2++) {
          double dt = 0.0;
          for (int i = tid; i < n; i += 64) dt += U[(long long)i * n + q2] * U[(long long)i * n + p];
          red[tid] = dt;
          __syncthreads();
          for (int o = 32; o > 0; o >>= 1) {
            if (tid < o) red[tid] += red[tid + o];
            __syncthreads();
          }
          dt = red[0];
          __syncthreads();
          for (int i = tid; i < n; i += 64) U[(long long)i * n + p] -= dt * U[(long long)i * n + q2];
          __syncthreads();
        }
        double nr = 0.0;
        for (int i = tid; i < n; i += 64) {
          const double v = U[(long long)i * n + p];
          nr += v * v;
        }
        red[tid] = nr;
        __syncthreads();
        for (int o = 32; o > 0; o >>= 1) {
          if (tid < o) red[tid] += red[tid + o];
          __syncthreads();
        }
        const double inv = red[0] > 0.0 ? 1.0 / sqrt(red[0]) : 0.0;
        __syncthreads();
        for (int i = tid; i < n; i += 64) U[(long long)i * n + p] *= inv;
        __syncthreads();
      }
    }
  }
}

// The WY factor of panel p of each block: T_p upper triangular with H_{j0} .. H_{j0+nbp-1} =
// I - V_p T_p V_p^T (dlarft, forward and columnwise), and V_p T_p written over the block's working
// matrix (columns j0 .. j0 + nbp - 1, rows j0 + 1 .. n - 1).
__global__ void __launch_bounds__(256) k_ht_larft(const HtBlock* __restrict__ bl, int nblk, const double* __restrict__ Va,
                                                  const double* __restrict__ Wa, double* __restrict__ Aa) {
  int g = blockIdx.x, b = -1;
  for (int i = 0; i < nblk; i++) {
    const int P = (bl[i].n - 1 + HT_NB - 1) / HT_NB;
    if (g < P) {
      b = i;
      break;
    }
    g -= P;
  }
  if (b < 0) return;
  const HtBlock B = bl[b];
  const int n = B.n, j0 = g * HT_NB, nbp = min(HT_NB, n - 1 - j0);
  const HtWs w = ht_ws(const_cast<double*>(Wa) + B.w_off, n, B.T);
  const double* V = Va + B.a_off;
  double* Y = Aa + B.a_off;
  __shared__ double G[HT_NB][HT_NB + 1], Tm[HT_NB][HT_NB + 1], St[64][HT_NB + 1];
  const int tid = threadIdx.x, ga = tid >> 3, gb = tid & 7;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int r0 = j0 + 1; r0 < n; r0 += 64) {
    for (int e2 = tid; e2 < 64 * HT_NB; e2 += 256) {
      const int l = e2 / HT_NB, j = e2 % HT_NB, r = r0 + l;
      St[l][j] = r < n && j < nbp ? V[(long long)r * n + j0 + j] : 0.0;
    }
    __syncthreads();
    for (int l = 0; l < 64; l++) {
      const double x = St[l][ga];
#pragma unroll
      for (int q = 0; q < 4; q++) acc[q] += x * St[l][gb + 8 * q];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 4; q++) G[ga][gb + 8 * q] = acc[q];
  if (tid < HT_NB)
    for (int q = 0; q < HT_NB; q++) Tm[tid][q] = 0.0;
  __syncthreads();
  for (int i = 0; i < nbp; i++) {
    const double ti = w.tau[j0 + i];
    double x = 0.0;
    if (tid < i)
      for (int l = tid; l < i; l++) x += Tm[tid][l] * G[l][i];
    __syncthreads();
    if (tid < i) Tm[tid][i] = -ti * x;
    if (tid == i) Tm[i][i] = ti;
    __syncthreads();
  }
  // V_p T_p rows j0+1 .. n-1
  for (int e2 = tid; e2 < (n - j0 - 1) * HT_NB; e2 += 256) {
    const int r = j0 + 1 + e2 / HT_NB, jj = e2 % HT_NB;
    if (jj >= nbp) continue;
    double x = 0.0;
    for (int i = 0; i <= jj; i++) x += V[(long long)r * n + j0 + i] * Tm[i][jj];
    Y[(long long)r * n + j0 + jj] = x;
  }
}

// ----------------------------------------------------------------------------- launchers
void launch_ht_first(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* A, double* W, hipStream_t s) {
  int g = 0;
  for (const auto& b : bl) g += b.T;
  if (g) NRGK_LAUNCH(k_ht_first, dim3(g), dim3(64), 0, s, d_bl, (int)bl.size(), A, W);
  NRGK_HIP(hipGetLastError());
}
int launch_ht_symv(const HtBlock* d_bl, const std::vector<HtBlock>& bl, int k, const double* A, const double* V, double* W,
                   hipStream_t s) {
  const int g = ht_grid(bl, k, 0);
  if (g) NRGK_LAUNCH(k_ht_symv, dim3(g), dim3(256), 0, s, d_bl, (int)bl.size(), k, A, V, W);
  NRGK_HIP(hipGetLastError());
  return g;
}
int launch_ht_update(const HtBlock* d_bl, const std::vector<HtBlock>& bl, int k, const double* A, double* V, double* W,
                     hipStream_t s) {
  const int g = ht_grid(bl, k, 1);
  if (g) NRGK_LAUNCH(k_ht_update, dim3(g), dim3(256), 0, s, d_bl, (int)bl.size(), k, A, V, W);
  NRGK_HIP(hipGetLastError());
  return g;
}
int launch_ht_syr2k(const HtBlock* d_bl, const std::vector<HtBlock>& bl, int k, double* A, const double* V, double* W,
                    hipStream_t s) {
  const int g = ht_grid(bl, k, 2);
  if (g) NRGK_LAUNCH(k_ht_syr2k, dim3(g), dim3(256), 0, s, d_bl, (int)bl.size(), k, A, V, W);
  NRGK_HIP(hipGetLastError());
  return g;
}
void launch_ht_stebz(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* W, double* lam, hipStream_t s) {
  int g = 0, nmax = 1;
  for (const auto& b : bl) {
    g += (b.n + 255) / 256;
    nmax = std::max(nmax, b.n);
  }
  if (g) NRGK_LAUNCH(k_ht_stebz, dim3(g), dim3(256), (size_t)nmax * 2 * sizeof(double), s, d_bl, (int)bl.size(), W, lam);
  NRGK_HIP(hipGetLastError());
}
void launch_ht_stein(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* W, const double* lam, double* U,
                     double* S, double ctol_rel, int* queue, hipStream_t s) {
  int g = 0, nmax = 1;
  for (const auto& b : bl) {
    g += b.mpad / 256;
    nmax = std::max(nmax, b.n);
  }
  if (!g) return;
  NRGK_HIP(hipMemsetAsync(queue, 0, sizeof(int), s));
  NRGK_LAUNCH(k_ht_stein, dim3(g), dim3(256), (size_t)nmax * 2 * sizeof(double), s, d_bl, (int)bl.size(), W, lam, U, S,
              ctol_rel, queue);
  NRGK_HIP(hipGetLastError());
  // the large clusters: a fixed grid that walks the queue (its length is read on the device)
  NRGK_LAUNCH(k_ht_stein_wide, dim3(1024), dim3(64), 0, s, d_bl, W, lam, U, S, ctol_rel, (const int*)queue);
  NRGK_HIP(hipGetLastError());
}
void launch_ht_larft(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* V, const double* W, double* A,
                     hipStream_t s) {
  int g = 0;
  for (const auto& b : bl) g += (b.n - 1 + HT_NB - 1) / HT_NB;
  if (g) NRGK_LAUNCH(k_ht_larft, dim3(g), dim3(256), 0, s, d_bl, (int)bl.size(), V, W, A);
  NRGK_HIP(hipGetLastError());
}
size_t ht_lds_bytes(int nmax) { return (size_t)nmax * 2 * sizeof(double); }
void set_ht_lds_limits(int bytes) {
  NRGK_HIP(hipFuncSetAttribute((const void*)k_ht_stebz, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  NRGK_HIP(hipFuncSetAttribute((const void*)k_ht_stein, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
}

}  // namespace nrgk
