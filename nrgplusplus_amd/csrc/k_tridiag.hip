// K3-K5: the eigensolver of the wide sector blocks (above the batched Jacobi's crossover), replacing
// per-block LAPACKE_dsyevd (nrgcore/include/utils/qmatrix.hpp:375-404) for those blocks.
//
// 1. Blocked Householder tridiagonalisation in FP64 (the lower-triangle form: reflector H_k = I -
//    tau_k v_k v_k^T annihilates column k below its subdiagonal).  Columns are reduced in panels of
//    HT_NB; inside a panel the trailing matrix is not touched: each column's symmetric product
//    y = A v is formed from the panel-start matrix and corrected by the panel's V and W, and the
//    panel ends with one rank-2k update A -= V W^T + W V^T.  All wide blocks of an iteration advance
//    column by column in the same launches:
//      k_ht_symv    one workgroup per 64 x 64 tile of the trailing lower triangle (the working
//                   matrix holds its lower tiles contiguously, see ht_at): the tile's row sums (A v)
//                   and column sums (A^T v over the strict lower part), v^T A v and the panel's
//                   V^T v, W^T v shares -- the HBM-bound bulk (each element read once per column,
//                   the lower triangle only)
//      k_ht_update  one workgroup per 64-row tile-row: reduces those partials to w, stores v and w,
//                   and forms the next column with the panel corrections and its norm
//      k_ht_syr2k   the panel's rank-2k update of the trailing lower triangle; the tiles of the
//                   next column also extract it
//    Every scalar a later launch needs (the reflector of column k from the norm partials) is
//    recomputed by each workgroup from the same partials in the same order, so all agree bit for bit
//    and no launch waits on a single-workgroup reduction.
// 2. Eigenvalues of T by bisection on Sturm counts (one thread per eigenvalue, T in LDS), ascending
//    as dsyevd returns them.
// 3. Eigenvectors of T by inverse iteration with partial-pivoting LU of T - lambda I (one lane per
//    eigenvector, all iterated at once; after every step the members of each cluster of eigenvalues
//    closer than ctol are made orthonormal by modified Gram-Schmidt, one workgroup per cluster from
//    a cluster list built on the device), for the columns that are needed: all of them, or only the
//    kept ones.
// 4. Back-transformation U = Q Z in WY form, Q_p = I - V_p T_p V_p^T per panel (k_ht_larft forms
//    V_p T_p): per panel Y = V_p^T Z and Z -= (V_p T_p) Y on the FP64 matrix cores (k_ht_wy_y,
//    k_ht_wy_z).
#include <cmath>

#include <hip/hip_ext.h>

#include "common.hpp"
#include "kernels.hpp"

namespace nrgk {

typedef double dbl4 __attribute__((ext_vector_type(4)));

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
// A wide block's working matrix holds its lower 64 x 64 tiles one after the other, tile (I, J) at
// 4096 ht_tidx(I, J), each row-major: a tile's 32 KB are contiguous (the symmetric products read
// ~10% faster than from 512-byte row pieces of a row-major matrix).  Entry (r, c), c <= r:
__device__ __forceinline__ long long ht_at(int r, int c) {
  return ht_tidx(r >> 6, c >> 6) * 4096 + (r & 63) * 64 + (c & 63);
}
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
// whose panel ends at k.  The launch carries the running task counts of the first HT_LMAX blocks in
// its arguments (end[i] = tasks of blocks 0..i), so the lookup is a scan over kernel-argument words;
// with more blocks (nb < 0) the counts are formed from the block table.
constexpr int HT_LMAX = 64;
struct HtLaunch {
  int nb;  // blocks taking part (a prefix of the table), or -1: scan the table
  int end[HT_LMAX];
};
template <int KIND>
__device__ __forceinline__ bool ht_find(const HtBlock* bl, int nblk, const HtLaunch& L, int k, int g, int& b,
                                        int& t) {
  if (L.nb >= 0) {
    int prev = 0;
    for (int i = 0; i < L.nb; i++) {
      if (g < L.end[i]) {
        b = i;
        t = g - prev;
        return true;
      }
      prev = L.end[i];
    }
    return false;
  }
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
// the same counts on the host: the grid size and the launch's table of running counts
int ht_grid(const std::vector<HtBlock>& bl, int k, int kind) {
  int g = 0;
  for (const auto& b : bl) {
    if (k > b.n - 2) break;
    if (kind == 2 && !ht_panel_end(b.n, k)) continue;
    g += kind == 1 ? ht_rows(b.T, k + 1) : ht_tiles(b.T, k + 1);
  }
  return g;
}
static HtLaunch ht_launch_table(const std::vector<HtBlock>& bl, int k, int kind) {
  HtLaunch L{};
  L.nb = 0;
  int g = 0;
  for (const auto& b : bl) {
    if (k > b.n - 2) break;
    if (L.nb == HT_LMAX) {
      L.nb = -1;
      return L;
    }
    if (!(kind == 2 && !ht_panel_end(b.n, k))) g += kind == 1 ? ht_rows(b.T, k + 1) : ht_tiles(b.T, k + 1);
    L.end[L.nb++] = g;
  }
  return L;
}

// The reflector of column k (dlarfg's form): from alpha = A(k+1, k) and the squared norm of the
// entries below it, the sum of the tile-rows' partials.  One wavefront sums them (lane l takes
// tile-rows t0 + l, t0 + l + 64, ..., then a butterfly) and lane 0 writes the reflector: every
// workgroup that needs it runs the same reduction in the same order, so all agree bit for bit.
struct HtRefl {
  double alpha, beta, tau, sc;
};
__device__ void ht_reflector_wave(const HtWs& w, int n, int T, int k, int lane, HtRefl* out) {
  const int s = k + 1;
  const double* ss = w.ss + (long long)(k & 1) * T;
  const double alpha = lane == 0 ? w.u[(long long)(k & 1) * n + s] : 0.0;
  double x = 0.0;
  for (int I = s / HT_TB + lane; I < T; I += 64) x += ss[I];
  x = ht_wave_sum(x);
  if (lane == 0) {
    HtRefl r;
    r.alpha = alpha;
    if (x == 0.0) {
      r.beta = alpha;
      r.tau = 0.0;
      r.sc = 0.0;
    } else {
      const double nrm = sqrt(alpha * alpha + x);
      r.beta = alpha >= 0.0 ? -nrm : nrm;
      r.tau = (r.beta - alpha) / r.beta;
      r.sc = 1.0 / (alpha - r.beta);
    }
    *out = r;
  }
}
// entry r of v_k (v[k+1] = 1, zero above) from the column's entry ur = u[r]
__device__ __forceinline__ double ht_v_of(double ur, const HtRefl& R, int s, int n, int r) {
  return (r < s || r >= n) ? 0.0 : (r == s ? 1.0 : ur * R.sc);
}
__device__ __forceinline__ double ht_v(const double* u, const HtRefl& R, int s, int n, int r) {
  return (r < s || r >= n) ? 0.0 : (r == s ? 1.0 : u[r] * R.sc);
}

// Every block's H (row-major, n x n, from the assembled blocks) into its working matrix's lower tiles
// (entries beyond n zero): one workgroup per tile, a row of 64 entries per 64 lanes.
__global__ void __launch_bounds__(256) k_ht_tile_in(const HtBlock* __restrict__ bl, int nblk, const double* __restrict__ H,
                                                    double* __restrict__ Aa) {
  int g = blockIdx.x, b = -1;
  for (int i = 0; i < nblk; i++) {
    const int nt = bl[i].T * (bl[i].T + 1) / 2;
    if (g < nt) {
      b = i;
      break;
    }
    g -= nt;
  }
  if (b < 0) return;
  const HtBlock B = bl[b];
  const int n = B.n;
  int I, J;
  ht_tile_of(g, I, J);
  const double* src = H + B.a_off;
  double* dst = Aa + B.p_off + (long long)g * 4096;
  const int c = J * HT_TB + (threadIdx.x & 63);
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int rr = (threadIdx.x >> 6) * 16 + i, r = I * HT_TB + rr;
    dst[rr * 64 + (threadIdx.x & 63)] = r < n && c < n ? src[(long long)r * n + c] : 0.0;
  }
}
void launch_ht_tile_in(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* H, double* A, hipStream_t s) {
  int g = 0;
  for (const auto& b : bl) g += b.T * (b.T + 1) / 2;
  if (g) NRGK_LAUNCH(k_ht_tile_in, dim3(g), dim3(256), 0, s, d_bl, (int)bl.size(), H, A);
  NRGK_HIP(hipGetLastError());
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
  const double* A = Aa + B.p_off;
  const int r = g * HT_TB + threadIdx.x;
  double sq = 0.0;
  if (r < n) {
    const double a = A[ht_at(r, 0)];
    if (r == 0) w.d[0] = a;
    else w.u[r] = a;
    if (r >= 2) sq = a * a;
  }
  sq = ht_wave_sum(sq);
  if (threadIdx.x == 0) w.ss[g] = sq;  // column 0: buffer 0
}

// One reduce-scatter exchange over a[0 .. 2H): the lane keeps a[H ..) if hi, else a[0 .. H), sends the
// other half to its DPP partner (CTRL) and adds what the partner sends into a[0 .. H).  The partner
// holds the same rows with hi the other way round, and every lane of the wavefront takes part.
template <int H, int CTRL>
__device__ __forceinline__ void ht_rs_step(double* a, bool hi) {
#pragma unroll
  for (int i = 0; i < H; i++) {
    const double send = hi ? a[i] : a[i + H];
    const double keep = hi ? a[i + H] : a[i];
    a[i] = keep + __builtin_amdgcn_update_dpp(0.0, send, CTRL, 0xF, 0xF, false);
  }
}

__global__ void __launch_bounds__(256) k_ht_symv(const HtBlock* __restrict__ bl, int nblk, const HtLaunch L, int k,
                                                 const double* __restrict__ Aa, const double* __restrict__ Va,
                                                 double* __restrict__ Wa) {
  int b, t;
  if (!ht_find<0>(bl, nblk, L, k, blockIdx.x, b, t)) return;
  const HtBlock B = bl[b];
  const int n = B.n, T = B.T, s = k + 1, t0 = s / HT_TB;
  const HtWs w = ht_ws(Wa + B.w_off, n, T);
  const double* A = Aa + B.p_off;
  const double* V = Va + B.p_off;
  const long long ld = B.ld;
  const double* u = w.u + (long long)(k & 1) * n;
  int Ip, Jp;
  ht_tile_of(t, Ip, Jp);
  const int I = t0 + Ip, J = t0 + Jp;
  __shared__ HtRefl R;
  __shared__ double vr[HT_TB], vc[HT_TB];
  __shared__ double red[4][4][16];
  __shared__ double cred[4][HT_TB];
  __shared__ double sp2[2];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  // everything the tile reads is requested at once: its entries (wave wv: rows 16 wv .. 16 wv + 15,
  // lane: column, one 512-byte row segment per load), the column under reduction at its rows (wave
  // 0) and columns (wave 1), and the reflector's norm partials (wave 2)
  const int c = J * HT_TB + lane;
  const bool diag = I == J;
  double a[16];
  const int r0 = I * HT_TB + __builtin_amdgcn_readfirstlane(wv) * 16;  // the same in every lane
  const double* At = A + ht_tidx(I, J) * 4096 + (r0 - I * HT_TB) * 64 + lane;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = r0 + i;
    const bool ok = r < n && c < n && r >= s && c >= s && (!diag || c <= r);
    a[i] = ok ? At[i * 64] : 0.0;
  }
  const int rv = (wv == 0 ? I : J) * HT_TB + lane;
  const double uv = wv < 2 && rv < n ? u[rv] : 0.0;
  if (wv == 2) ht_reflector_wave(w, n, T, k, lane, &R);
  __syncthreads();
  if (wv == 0) vr[lane] = ht_v_of(uv, R, s, n, rv);
  else if (wv == 1) vc[lane] = ht_v_of(uv, R, s, n, rv);
  __syncthreads();
  double colacc = 0.0;
  const double vcl = vc[lane];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = r0 + i;
    colacc += (!diag || c < r) ? a[i] * vr[r - I * HT_TB] : 0.0;
    a[i] *= vcl;
  }
  cred[wv][lane] = colacc;
  // row sums of the wave's 16 x 64 products over each 16-lane row of the wavefront, reduce-scattered
  // in registers: four exchanges with the partner lane p ^ 15, p ^ 7, p ^ 1, p ^ 2 (DPP row_mirror,
  // row_half_mirror, quad_perm), each keeping the half of the rows the lane goes on with, leave lane
  // p with row 8 p3 + 4 p2 + 2 p0 + p1 (p = lane & 15, bits p3..p0) summed over its 16 lanes
  const int p = lane & 15;
  ht_rs_step<8, 0x140>(a, (p & 8) != 0);
  ht_rs_step<4, 0x141>(a, (p & 4) != 0);
  ht_rs_step<2, 0xB1>(a, (p & 1) != 0);
  ht_rs_step<1, 0x4E>(a, (p & 2) != 0);
  red[wv][lane >> 4][((p >> 3) << 3) | (((p >> 2) & 1) << 2) | ((p & 1) << 1) | ((p >> 1) & 1)] = a[0];
  __syncthreads();
  const long long ix = ht_tidx(I, J);
  double myv = 0.0;
  if (tid < 64) {
    const int h = tid >> 4, i = tid & 15;
    const double rs = (red[h][0][i] + red[h][1][i]) + (red[h][2][i] + red[h][3][i]);
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
  if (tid == 0) w.sp[t] = sp2[0] + sp2[1];  // by tile index within the launch
  if (diag) {
    // this tile-row's shares of V^T v and W^T v over the panel's earlier columns: wave h sums rows
    // 16 h .. 16 h + 15 (lane: V or W, column j; all sixteen loads at once), then the quarters in order
    const int j0 = (k / HT_NB) * HT_NB, cc = k - j0;
    const int arr = lane >> 5, j = lane & 31;
    double acc = 0.0;
    if (j < cc) {
      double x[16];
#pragma unroll
      for (int l = 0; l < 16; l++) {
        const int r = I * HT_TB + wv * 16 + l;
        x[l] = r >= s && r < n ? (arr == 0 ? V[r * ld + j0 + j] : w.W[(long long)r * HT_NB + j]) : 0.0;
      }
#pragma unroll
      for (int l = 0; l < 16; l++) acc += x[l] * vr[wv * 16 + l];
    }
    cred[wv][lane] = acc;
    __syncthreads();
    if (tid < 64 && j < cc) {
      const double q = (cred[0][tid] + cred[1][tid]) + (cred[2][tid] + cred[3][tid]);
      if (arr == 0) w.qv[I * HT_NB + j] = q;
      else w.qw[I * HT_NB + j] = q;
    }
  }
}

// Quarter h (0..3) of y(r) = (A v)(r) from the tile partials: the list of the row sums of tiles
// (I, t0..I) followed by the column sums of tiles (I..T-1, I), entries h, h + 4, ... of it.
__device__ __forceinline__ double ht_ypart(const HtWs& w, int T, int t0, int r, int h) {
  const int I = r / HT_TB, q = r - I * HT_TB, nr = I - t0 + 1, L = nr + T - I;
  double y = 0.0;
  for (int e0 = h; e0 < L; e0 += 4 * 8) {
    double x[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int e = e0 + 4 * j;
      x[j] = e >= L ? 0.0 : (e < nr ? w.rp[ht_tidx(I, t0 + e) * 64 + q] : w.cp[ht_tidx(I + e - nr, I) * 64 + q]);
    }
#pragma unroll
    for (int j = 0; j < 8; j++) y += x[j];
  }
  return y;
}
// V(r,:) W^T v + W(r,:) V^T v over the panel's cc earlier columns
__device__ __forceinline__ double ht_corr(const HtWs& w, const double* V, long long ld, int j0, int cc,
                                          const double* Vv, const double* Wv, int r) {
  double corr = 0.0;
#pragma unroll 8
  for (int j = 0; j < cc; j++) corr += V[r * ld + j0 + j] * Wv[j] + w.W[(long long)r * HT_NB + j] * Vv[j];
  return corr;
}

// One workgroup per 64-row tile-row of the trailing matrix: w(r) = tau (y(r) - V(r,:) W^T v -
// W(r,:) V^T v) - (tau/2) (w^T v) v(r) for its rows, v and w stored, then the next column with the
// panel corrections and its norm partial.  Threads 0..255 split each row's partial sums in quarters
// (thread 4 q + h: row q, quarter h); threads 256..259 do the same for row k + 1, whose w every
// workgroup needs, so its value is the same bits as the one its owner stores.
__global__ void __launch_bounds__(320) k_ht_update(const HtBlock* __restrict__ bl, int nblk, const HtLaunch L, int k,
                                                   const double* __restrict__ Aa, double* __restrict__ Va,
                                                   double* __restrict__ Wa) {
  int b, t;
  if (!ht_find<1>(bl, nblk, L, k, blockIdx.x, b, t)) return;
  const HtBlock B = bl[b];
  const int n = B.n, T = B.T, s = k + 1, t0 = s / HT_TB, I = t0 + t;
  const int j0 = (k / HT_NB) * HT_NB, cc = k - j0;
  const bool pend = ht_panel_end(n, k);
  const HtWs w = ht_ws(Wa + B.w_off, n, T);
  const double* A = Aa + B.p_off;
  double* V = Va + B.p_off;
  const long long ld = B.ld;
  const double* u = w.u + (long long)(k & 1) * n;
  __shared__ HtRefl R;
  __shared__ double Vv[HT_NB], Wv[HT_NB];
  __shared__ double qpart[4][2][HT_NB];
  __shared__ double sred[4];
  __shared__ double ypart[4][HT_TB], yspart[4];
  __shared__ double sh_wv, sh_w1;
  __shared__ double sh_ws[HT_NB], sh_vs[HT_NB];  // W(s, :) and V(s, j0 + :), the panel's row k + 1
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  if (wv == 4) ht_reflector_wave(w, n, T, k, lane, &R);
  // v^T A v over the launch's tiles (partials by tile index): a fixed partition and order, the same in
  // every workgroup
  const int m = T - t0, NTa = m * (m + 1) / 2;
  if (tid < 256) {
    double acc = 0.0;
    for (int q0 = tid; q0 < NTa; q0 += 256 * 8) {
      double x[8];
#pragma unroll
      for (int j = 0; j < 8; j++) x[j] = q0 + 256 * j < NTa ? w.sp[q0 + 256 * j] : 0.0;
#pragma unroll
      for (int j = 0; j < 8; j++) acc += x[j];
    }
    acc = ht_wave_sum(acc);
    if (lane == 0) sred[wv] = acc;
    // the panel's V^T v and W^T v: thread 4 p + h sums quarter h of the tile-rows of pair p
    const int pr = tid >> 2, h = tid & 3, arr = pr >> 5, j = pr & 31;
    if (j < cc) {
      const double* src = arr == 0 ? w.qv : w.qw;
      double a = 0.0;
      for (int i0 = t0 + h; i0 < T; i0 += 4 * 8) {
        double x[8];
#pragma unroll
        for (int jj = 0; jj < 8; jj++) {
          const int Ir = i0 + 4 * jj;
          x[jj] = Ir < T ? src[Ir * HT_NB + j] : 0.0;
        }
#pragma unroll
        for (int jj = 0; jj < 8; jj++) a += x[jj];
      }
      qpart[h][arr][j] = a;
    }
    // the partial sums of this workgroup's rows
    const int r = I * HT_TB + pr;
    ypart[h][pr] = r < n && r >= s ? ht_ypart(w, T, t0, r, h) : 0.0;
  } else if (tid < 260 && !pend) {
    yspart[tid - 256] = ht_ypart(w, T, t0, s, tid - 256);
  }
  __syncthreads();
  if (tid < 2 * HT_NB) {
    const int arr = tid >> 5, j = tid & 31;
    if (j < cc) {
      const double a = (qpart[0][arr][j] + qpart[1][arr][j]) + (qpart[2][arr][j] + qpart[3][arr][j]);
      if (arr == 0) Vv[j] = a;
      else Wv[j] = a;
    }
  } else if (tid < 2 * HT_NB + HT_NB && !pend) {
    const int j = tid - 2 * HT_NB;
    if (j < cc) {
      sh_ws[j] = w.W[(long long)s * HT_NB + j];
      sh_vs[j] = V[s * ld + j0 + j];
    }
  }
  __syncthreads();
  if (tid == 0) {
    double vw = 0.0;
    for (int j = 0; j < cc; j++) vw += Vv[j] * Wv[j];
    sh_wv = R.tau * ((sred[0] + sred[1]) + (sred[2] + sred[3]) - 2.0 * vw);
  }
  __syncthreads();
  const double wvs = sh_wv;
  const int r = I * HT_TB + tid;
  double wr = 0.0, vrr = 0.0;
  const bool own = tid < 64 && r < n && r >= s;
  if (own) {
    const double y = (ypart[0][tid] + ypart[1][tid]) + (ypart[2][tid] + ypart[3][tid]);
    vrr = ht_v(u, R, s, n, r);
    wr = R.tau * (y - ht_corr(w, V, ld, j0, cc, Vv, Wv, r)) - 0.5 * R.tau * wvs * vrr;
    w.W[(long long)r * HT_NB + cc] = wr;
    V[r * ld + k] = vrr;
  } else if (tid == 256 && !pend) {
    const double y = (yspart[0] + yspart[1]) + (yspart[2] + yspart[3]);
    sh_w1 = R.tau * (y - ht_corr(w, V, ld, j0, cc, Vv, Wv, s)) - 0.5 * R.tau * wvs * 1.0;
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
    double un = A[ht_at(r, s)];
#pragma unroll 8
    for (int j = 0; j < cc; j++) un -= V[r * ld + j0 + j] * sh_ws[j] + w.W[(long long)r * HT_NB + j] * sh_vs[j];
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
__global__ void __launch_bounds__(256) k_ht_syr2k(const HtBlock* __restrict__ bl, int nblk, const HtLaunch L, int k,
                                                  double* __restrict__ Aa, const double* __restrict__ Va,
                                                  double* __restrict__ Wa) {
  int b, t;
  if (!ht_find<2>(bl, nblk, L, k, blockIdx.x, b, t)) return;
  const HtBlock B = bl[b];
  const int n = B.n, T = B.T, s = k + 1, t0 = s / HT_TB;
  const int j0 = (k / HT_NB) * HT_NB, nbu = k - j0 + 1;
  const HtWs w = ht_ws(Wa + B.w_off, n, T);
  double* A = Aa + B.p_off;
  const double* V = Va + B.p_off;
  const long long ld = B.ld;
  int Ip, Jp;
  ht_tile_of(t, Ip, Jp);
  const int I = t0 + Ip, J = t0 + Jp;
  // thread: rows 4 tr .. 4 tr + 3, columns tc, tc + 16, tc + 32, tc + 48 of the tile (sixteen lanes
  // read sixteen consecutive entries of a staged row: no LDS bank conflicts)
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  __shared__ double VI[16][HT_TB + 1], WI[16][HT_TB + 1], VJ[16][HT_TB + 1], WJ[16][HT_TB + 1];
  __shared__ double sqp[16];
  const bool diag = I == J;
  // the tile's entries are requested before the product, so their loads overlap it
  double av[4][4], acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = I * HT_TB + 4 * tr + i;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int c = J * HT_TB + tc + 16 * j;
      const bool ok = r < n && c < n && r >= s && c >= s && (!diag || c <= r);
      av[i][j] = ok ? A[ht_at(r, c)] : 0.0;
      acc[i][j] = 0.0;
    }
  }
  for (int jb = 0; jb < nbu; jb += 16) {
    // stage 16 panel columns of V and W for the tile's rows (I) and columns (J): sixteen lanes read
    // 16 consecutive entries of one row
    for (int e = tid; e < 16 * HT_TB; e += 256) {
      const int jj = e & 15, l = e >> 4, j = jb + jj;
      const int ri = I * HT_TB + l, rj = J * HT_TB + l;
      const bool okj = j < nbu;
      VI[jj][l] = okj && ri < n ? V[ri * ld + j0 + j] : 0.0;
      WI[jj][l] = okj && ri < n ? w.W[(long long)ri * HT_NB + j] : 0.0;
      VJ[jj][l] = okj && rj < n ? V[rj * ld + j0 + j] : 0.0;
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
        vj[q] = VJ[jj][tc + 16 * q];
        wj[q] = WJ[jj][tc + 16 * q];
      }
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] += vi[i] * wj[j] + wi[i] * vj[j];
    }
    __syncthreads();
  }
  double sq = 0.0;
  const int cs = s - J * HT_TB;  // column s inside this tile (when J == t0)
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = I * HT_TB + 4 * tr + i;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int c = J * HT_TB + tc + 16 * j;
      if (r < n && c < n && r >= s && c >= s && (!diag || c <= r)) {
        const double na = av[i][j] - acc[i][j];
        A[ht_at(r, c)] = na;
        if (c == s) {
          if (r == s) w.d[s] = na;
          else w.u[(long long)(s & 1) * n + r] = na;
          if (r >= s + 2) sq += na * na;
        }
      }
    }
  }
  if (J == t0) {
    // the threads holding column s: tc == cs % 16 (one per 4 rows); the tile-row's norm partial
    if (tc == (cs & 15)) sqp[tr] = sq;
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
  if (tnorm == 0.0) {  // T = 0: every eigenvalue is exactly zero
    lam[B.e_off + j] = 0.0;
    return;
  }
  double a = rlo[0] - 2.0 * eps * tnorm - pivmin, c = rhi[0] + 2.0 * eps * tnorm + pivmin;
  for (int it = 0; it < 200; it++) {
    const double mid = 0.5 * (a + c);
    if (c - a <= 2.0 * eps * fmax(fabs(a), fabs(c)) + atol || mid <= a || mid >= c) break;
    // Sturm count: eigenvalues below mid.  e_{i-1}^2 / q by the hardware reciprocal and one Newton
    // step (|q| >= pivmin keeps it finite, and e^2 / pivmin below the overflow threshold), which
    // shortens the dependent chain of each step to three fused operations
    int cnt = 0;
    double q = d[0] - mid;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
    for (int i = 1; i < n; i++) {
      double r = __builtin_amdgcn_rcp(q);
      r = fma(r, fma(-q, r, 1.0), r);
      q = fma(-e2[i - 1], r, d[i] - mid);
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

// Rows per register tile of the column loops below: each tile's loads are issued together, so a
// lane waits for memory once per tile rather than once per row of its serial recurrences.
constexpr int HT_RT = 8;

// row i0 .. i0 + HT_RT - 1 of the forward elimination: the right-hand side of the next row and T's entries
__device__ __forceinline__ void ht_fwd_load(const double* __restrict__ d, const double* __restrict__ e, int n,
                                            const double* __restrict__ U, int p, int i0, double* xn, double* sb,
                                            double* dn, double* en) {
  // rows past the end read the last row instead (every load is issued, with no branch around it);
  // their values are never used
#pragma unroll
  for (int j = 0; j < HT_RT; j++) {
    const int i = min(i0 + j, n - 2);
    xn[j] = U[(long long)(i + 1) * n + p];
    sb[j] = e[i];
    dn[j] = d[i + 1];
    en[j] = e[i + 1];  // e[n - 1] is 0
  }
}
// rows i1, i1 - 1, .. of the back substitution: the eliminated right-hand side and the factors
__device__ __forceinline__ void ht_bwd_load(const double* __restrict__ U, int n, int p, int i1,
                                            const double* __restrict__ f0, const double* __restrict__ f1,
                                            const double* __restrict__ f2, long long mp, int slot, double* uu,
                                            double* a0, double* a1, double* a2) {
#pragma unroll
  for (int j = 0; j < HT_RT; j++) {
    const int i = max(i1 - j, 0);
    const long long o = (long long)i * mp + slot;
    uu[j] = U[(long long)i * n + p];
    a0[j] = f0[o];
    a1[j] = f1[o];
    a2[j] = f2[o];
  }
}
// One inverse-iteration step for eigenvalue lp: the iterate in column p of U (stride n) is replaced by
// (T - lp I)^-1 times it.  T - lp I = P L U with partial pivoting; the factorisation and the
// elimination of the iterate share one pass (the L factor is not kept), the upper factor's three
// diagonals go to the scratch f0 (as the pivot's reciprocal), f1, f2 at entry i * mp + slot, then one
// backward pass.  A pivot smaller than piv0 in magnitude is moved out to +-piv0 (dlagtf/dlagts'
// perturbation), which bounds the growth of one step by about 1 / piv0.
// Each pass requests the next HT_RT rows before it stores the current ones: a wait on vector memory
// also waits for every store issued before the awaited load, so the loads are not held behind them.
// the forward elimination over rows i0 .. i0 + HT_RT - 1 (cd, cu: the pivot row's diagonal and
// superdiagonal, xi: its right-hand side)
__device__ __forceinline__ void ht_fwd_rows(int n, double lp, double piv0, double* __restrict__ U, int p,
                                            double* __restrict__ f0, double* __restrict__ f1,
                                            double* __restrict__ f2, long long mp, int slot, int i0,
                                            const double* xn, const double* sb, const double* dn, const double* en,
                                            double& cd, double& cu, double& xi) {
#pragma unroll
  for (int j = 0; j < HT_RT; j++) {
    const int i = i0 + j;
    if (i >= n - 1) continue;
    const double sub = sb[j], nd = dn[j] - lp, nu = en[j];
    const long long o = (long long)i * mp + slot;
    if (fabs(cd) >= fabs(sub)) {
      if (fabs(cd) < piv0) cd = cd < 0.0 ? -piv0 : piv0;
      const double r = 1.0 / cd, ml = sub * r;
      f0[o] = r;
      f1[o] = cu;
      f2[o] = 0.0;
      U[(long long)i * n + p] = xi;
      xi = xn[j] - ml * xi;
      cd = nd - ml * cu;
      cu = nu;
    } else {  // rows i and i + 1 exchanged
      const double r = 1.0 / sub, ml = cd * r;
      f0[o] = r;
      f1[o] = nd;
      f2[o] = nu;
      U[(long long)i * n + p] = xn[j];
      xi = xi - ml * xn[j];
      cd = cu - ml * nd;
      cu = -ml * nu;
    }
  }
}
// the back substitution over rows i1, i1 - 1, .. i1 - HT_RT + 1 (x1, x2: the solution's next two rows)
__device__ __forceinline__ void ht_bwd_rows(int n, double* __restrict__ U, int p, int i1, const double* uu,
                                            const double* a0, const double* a1, const double* a2, double& x1,
                                            double& x2) {
#pragma unroll
  for (int j = 0; j < HT_RT; j++) {
    const int i = i1 - j;
    if (i < 0) continue;
    const double v = (uu[j] - a1[j] * x1 - a2[j] * x2) * a0[j];
    U[(long long)i * n + p] = v;
    x2 = x1;
    x1 = v;
  }
}
__device__ void ht_inverse_step(const double* __restrict__ d, const double* __restrict__ e, int n, double lp,
                                double piv0, double* __restrict__ U, int p, double* __restrict__ f0,
                                double* __restrict__ f1, double* __restrict__ f2, long long mp, int slot) {
  double cd = d[0] - lp, cu = n > 1 ? e[0] : 0.0;
  double xi = U[p];
  if (n >= 2) {
    // two register tiles in turn: each is refilled while the other one's rows are eliminated
    double xa[HT_RT], sa[HT_RT], da[HT_RT], ea[HT_RT], xb[HT_RT], sb[HT_RT], db[HT_RT], eb[HT_RT];
    ht_fwd_load(d, e, n, U, p, 0, xa, sa, da, ea);
    for (int i0 = 0; i0 < n - 1; i0 += 2 * HT_RT) {
      ht_fwd_load(d, e, n, U, p, i0 + HT_RT, xb, sb, db, eb);
      __builtin_amdgcn_sched_barrier(0);  // keep these loads ahead of the stores below
      ht_fwd_rows(n, lp, piv0, U, p, f0, f1, f2, mp, slot, i0, xa, sa, da, ea, cd, cu, xi);
      ht_fwd_load(d, e, n, U, p, i0 + 2 * HT_RT, xa, sa, da, ea);
      __builtin_amdgcn_sched_barrier(0);
      ht_fwd_rows(n, lp, piv0, U, p, f0, f1, f2, mp, slot, i0 + HT_RT, xb, sb, db, eb, cd, cu, xi);
    }
  }
  if (fabs(cd) < piv0) cd = cd < 0.0 ? -piv0 : piv0;
  const long long ol = (long long)(n - 1) * mp + slot;
  f0[ol] = 1.0 / cd;
  f1[ol] = 0.0;
  f2[ol] = 0.0;
  U[(long long)(n - 1) * n + p] = xi;
  double x1 = 0.0, x2 = 0.0;
  double ua[HT_RT], a0[HT_RT], a1[HT_RT], a2[HT_RT], ub[HT_RT], b0[HT_RT], b1[HT_RT], b2[HT_RT];
  ht_bwd_load(U, n, p, n - 1, f0, f1, f2, mp, slot, ua, a0, a1, a2);
  for (int i1 = n - 1; i1 >= 0; i1 -= 2 * HT_RT) {
    ht_bwd_load(U, n, p, i1 - HT_RT, f0, f1, f2, mp, slot, ub, b0, b1, b2);
    __builtin_amdgcn_sched_barrier(0);
    ht_bwd_rows(n, U, p, i1, ua, a0, a1, a2, x1, x2);
    ht_bwd_load(U, n, p, i1 - 2 * HT_RT, f0, f1, f2, mp, slot, ua, a0, a1, a2);
    __builtin_amdgcn_sched_barrier(0);
    ht_bwd_rows(n, U, p, i1 - HT_RT, ub, b0, b1, b2, x1, x2);
  }
}

// inverse-iteration steps per eigenvector: from a random start one step leaves the vector off by about
// eps ||T|| / gap along its neighbours (below 5e-12 outside the clusters), the second polishes it
constexpr int HT_ITERS = 2;

// the infinity norm of T (d, e in memory p)
__device__ __forceinline__ double ht_tnorm_entry(const double* d, const double* e, int n, int i) {
  return fabs(d[i]) + (i < n - 1 ? fabs(e[i]) : 0.0) + (i > 0 ? fabs(e[i - 1]) : 0.0);
}

// first index >= x that starts a cluster (lambda more than ctol above its predecessor), or n
__device__ __forceinline__ int ht_cluster_start_at(const double* L, int n, int x, double ctol) {
  while (x > 0 && x < n && L[x] - L[x - 1] <= ctol) x++;
  return x;
}

// column p normalised, scaled by its largest entry first so that the squares of an iterate of any
// size stay finite
__device__ void ht_col_normalise(double* __restrict__ U, int n, int p) {
  // one read-only pass forms the largest magnitude and the plain sum of squares (sixteen rows a
  // batch); only when that sum over- or underflows does a second pass form it from the scaled
  // entries; then the column is scaled
  constexpr int RB = 2 * HT_RT;
  double mx = 0.0, s2 = 0.0;
  for (int i0 = 0; i0 < n; i0 += RB) {
    double a[RB];
#pragma unroll
    for (int j = 0; j < RB; j++) a[j] = i0 + j < n ? U[(long long)(i0 + j) * n + p] : 0.0;
#pragma unroll
    for (int j = 0; j < RB; j++) {
      mx = fmax(mx, fabs(a[j]));
      s2 += a[j] * a[j];
    }
  }
  if (mx == 0.0) return;
  double inv;
  if (s2 > 1e-280 && s2 < 1e280) {
    inv = 1.0 / sqrt(s2);
  } else {
    const double sc = 1.0 / mx;
    double nr = 0.0;
    for (int i0 = 0; i0 < n; i0 += RB) {
      double a[RB];
#pragma unroll
      for (int j = 0; j < RB; j++) a[j] = i0 + j < n ? U[(long long)(i0 + j) * n + p] * sc : 0.0;
#pragma unroll
      for (int j = 0; j < RB; j++) nr += a[j] * a[j];
    }
    inv = sc / sqrt(nr);
  }
  for (int i0 = 0; i0 < n; i0 += HT_RT) {
    double a[HT_RT];
#pragma unroll
    for (int j = 0; j < HT_RT; j++) a[j] = i0 + j < n ? U[(long long)(i0 + j) * n + p] : 0.0;
#pragma unroll
    for (int j = 0; j < HT_RT; j++)
      if (i0 + j < n) U[(long long)(i0 + j) * n + p] = a[j] * inv;
  }
}

// Eigenvectors of T for the eigenvalues 0 .. m-1 of each block (and the rest of a cluster that
// starts below m), into the columns of U.  Clusters are runs of eigenvalues each within ctol of the
// one before.  The host runs HT_ITERS rounds of
//   k_ht_stein_step  one inverse-iteration step for every wanted eigenvector (normalised), from a fixed
//                    pseudo-random start in the first round
//   k_ht_stein_mgs   every cluster of more than one member made orthonormal
// and one more k_ht_stein_mgs at the end.  Orthogonalising after every step matters: the members of a
// numerically degenerate cluster (copies of an eigenvalue in different split blocks of T) are all
// pulled toward the copy their shifts sit closest to, so steps without it leave them nearly
// dependent, and their orthogonalisation would then blow up the components outside the cluster;
// kept apart each step, those stay at eps ||T|| / gap.
//
// k_ht_stein_step: workgroup (one wavefront) g of a block takes the clusters that start in
// [HT_CHUNK g, HT_CHUNK (g + 1)) below m; lane t iterates the members a + t, a + t + 64, ... of that
// range (a = its first cluster start), with the LU factors in its scratch slot g * 64 + t of
// mpad = ht_stein_slots(m).
constexpr int HT_CHUNK = 48;
__device__ __forceinline__ double ht_tnorm_wave(const HtWs& w, int n, int lane) {
  double nm = 0.0;
  for (int i = lane; i < n; i += 64) nm = fmax(nm, ht_tnorm_entry(w.d, w.e, n, i));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) nm = fmax(nm, __shfl_xor(nm, o, 64));
  return nm;
}
__global__ void __launch_bounds__(64) k_ht_stein_step(const HtBlock* __restrict__ bl, int nblk,
                                                      const double* __restrict__ Wa, const double* __restrict__ lam,
                                                      double* __restrict__ Ua, double* __restrict__ Sa, double ctol_rel,
                                                      int first) {
  int g = blockIdx.x, b = -1;
  for (int i = 0; i < nblk; i++) {
    const int cw = (bl[i].m + HT_CHUNK - 1) / HT_CHUNK;
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
  const int lane = threadIdx.x;
  const double tnorm = ht_tnorm_wave(w, n, lane);
  const double* L = lam + B.e_off;
  const double ctol = ctol_rel * tnorm;
  const int a = ht_cluster_start_at(L, n, g * HT_CHUNK, ctol);
  const int bend = ht_cluster_start_at(L, n, min((g + 1) * HT_CHUNK, B.m), ctol);
  double* U = Ua + B.a_off;
  const int slot = g * 64 + lane;
  double* f0 = Sa + B.s_off;
  double* f1 = f0 + (long long)n * mp;
  double* f2 = f1 + (long long)n * mp;
  // pivots are kept at least eps ||T|| in magnitude (1 for T = 0, where any basis is one)
  const double piv0 = tnorm > 0.0 ? 2.220446049250313e-16 * tnorm : 1.0;
  // T in LDS: its entries are the same for every lane, and LDS reads do not wait behind the
  // elimination's stores to memory
  extern __shared__ double tlds[];
  double* td = tlds;
  double* te = tlds + n;
  for (int i = lane; i < n; i += 64) {
    td[i] = w.d[i];
    te[i] = i < n - 1 ? w.e[i] : 0.0;
  }
  __syncthreads();
  for (int p = a + lane; p < bend; p += 64) {
    if (first)
      for (int i = 0; i < n; i++) U[(long long)i * n + p] = ht_start((unsigned)p, (unsigned)i);
    ht_inverse_step(td, te, n, L[p], piv0, U, p, f0, f1, f2, mp, slot);
    ht_col_normalise(U, n, p);
  }
}

// The clusters of more than one member that start below m, one workgroup per block: (block, first
// member, members) triples appended to list, their number in count (zeroed beforehand).
__global__ void __launch_bounds__(256) k_ht_stein_clusters(const HtBlock* __restrict__ bl, const double* __restrict__ Wa,
                                                           const double* __restrict__ lam, double ctol_rel,
                                                           int* __restrict__ list, int* __restrict__ count) {
  const int b = blockIdx.x;
  const HtBlock B = bl[b];
  const int n = B.n, tid = threadIdx.x;
  const HtWs w = ht_ws(const_cast<double*>(Wa) + B.w_off, n, B.T);
  __shared__ double red[4];
  const double nm = ht_tnorm_wave(w, n, tid & 63);
  if ((tid & 63) == 0) red[tid >> 6] = nm;
  __syncthreads();
  const double tnorm = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const double ctol = ctol_rel * tnorm;
  const double* L = lam + B.e_off;
  for (int j = tid; j < B.m; j += 256) {
    const bool start = j == 0 || L[j] - L[j - 1] > ctol;
    if (!start || j + 1 >= n || L[j + 1] - L[j] > ctol) continue;
    int e = j + 1;
    while (e < n && L[e] - L[e - 1] <= ctol) e++;
    const int q = atomicAdd(count, 1);
    list[3 * q] = b;
    list[3 * q + 1] = j;
    list[3 * q + 2] = e - j;
  }
}

// Modified Gram-Schmidt over each listed cluster (columns cs .. cs + c - 1 of its block's U, stride
// n) by one 256-thread workgroup, rows split over the threads (thread t: rows t, t + 256, ...; each
// thread only ever touches its own rows), the workgroups taking the list in strides of the grid.
// Round r, for the later columns sixteen at a time: one sweep forms their dot products d_s with
// column cs + r (each thread's rows in order, a butterfly per wavefront, the four wavefronts' sums
// added in order, so every thread holds the same d), a second takes (d_s / d_r) column r out of
// them; column r is then scaled by 1 / sqrt(d_r).
constexpr int HT_MG = 16;
__global__ void __launch_bounds__(256) k_ht_stein_mgs(const HtBlock* __restrict__ bl, const int* __restrict__ list,
                                                      const int* __restrict__ count, double* __restrict__ Ua) {
  __shared__ double red[4][HT_MG];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int nq = *count;
  for (int q = blockIdx.x; q < nq; q += gridDim.x) {
    const HtBlock B = bl[list[3 * q]];
    const int cs = list[3 * q + 1], c = list[3 * q + 2], n = B.n;
    double* U = Ua + B.a_off;
    for (int r = 0; r < c; r++) {
      double drr = 0.0;
      for (int s0 = r; s0 < c; s0 += HT_MG) {
        double acc[HT_MG];
#pragma unroll
        for (int j = 0; j < HT_MG; j++) acc[j] = 0.0;
        for (int i = tid; i < n; i += 256) {
          const double* row = U + (long long)i * n + cs;
          const double xr = row[r];
#pragma unroll
          for (int j = 0; j < HT_MG; j++)
            if (s0 + j < c) acc[j] += xr * row[s0 + j];
        }
#pragma unroll
        for (int j = 0; j < HT_MG; j++) {
          const double v = ht_wave_sum(acc[j]);
          if (lane == 0) red[wv][j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < HT_MG; j++) acc[j] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
        __syncthreads();
        if (s0 == r) drr = acc[0];
        double coef[HT_MG];
#pragma unroll
        for (int j = 0; j < HT_MG; j++) coef[j] = s0 + j > r && s0 + j < c && drr > 0.0 ? acc[j] / drr : 0.0;
        for (int i = tid; i < n; i += 256) {
          double* row = U + (long long)i * n + cs;
          const double xr = row[r];
#pragma unroll
          for (int j = 0; j < HT_MG; j++)
            if (s0 + j > r && s0 + j < c) row[s0 + j] -= coef[j] * xr;
        }
      }
      const double inv = drr > 0.0 ? 1.0 / sqrt(drr) : 0.0;
      for (int i = tid; i < n; i += 256) U[(long long)i * n + cs + r] *= inv;
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
  const double* V = Va + B.p_off;
  double* Y = Aa + B.p_off;
  const long long ld = B.ld;
  __shared__ double G[HT_NB][HT_NB + 1], Tm[HT_NB][HT_NB + 1], St[64][HT_NB + 1];
  const int tid = threadIdx.x, ga = tid >> 3, gb = tid & 7;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int r0 = j0 + 1; r0 < n; r0 += 64) {
    for (int e2 = tid; e2 < 64 * HT_NB; e2 += 256) {
      const int l = e2 / HT_NB, j = e2 % HT_NB, r = r0 + l;
      St[l][j] = r < n && j < nbp ? V[r * ld + j0 + j] : 0.0;
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
    for (int i = 0; i <= jj; i++) x += V[r * ld + j0 + i] * Tm[i][jj];
    Y[r * ld + j0 + jj] = x;
  }
}

// The back-transformation U = Q Z of the wanted columns, one panel at a time (last first): with
// V_p (rows j0+1 .. n-1, columns j0 .. j0+nbp-1 of the Householder vectors) and V_p T_p (the same
// place in the working matrix, k_ht_larft), Y = V_p^T Z and Z -= (V_p T_p) Y.  The wanted columns
// are few in an RG step (tens per block), so the products run over rows instead of output tiles:
//   k_ht_wy_y  workgroup (block, 256-row chunk c, 64-column tile): the chunk's share of Y, a 32 x 64
//              partial, on the matrix cores with operands read straight from memory
//   k_ht_wy_z  the same workgroups: Y's tile summed over the chunks in order (into LDS), then the
//              chunk's 256 x 64 piece of Z updated on the matrix cores
// Each partial lives in the block's scratch at (c * 32 + j) * m + column.
constexpr int WY_ROWS = 256;
__device__ __forceinline__ bool ht_wy_find(const HtBlock* bl, int nblk, const HtLaunch& L, int step, int g, int& b,
                                           int& t) {
  if (L.nb >= 0) {
    int prev = 0;
    for (int i = 0; i < L.nb; i++) {
      if (g < L.end[i]) {
        b = i;
        t = g - prev;
        return true;
      }
      prev = L.end[i];
    }
    return false;
  }
  for (int i = 0; i < nblk; i++) {
    const int n = bl[i].n, P = (n - 1 + HT_NB - 1) / HT_NB;
    if (step >= P) continue;
    const int R = n - (P - 1 - step) * HT_NB - 1;
    const int cnt = ((R + WY_ROWS - 1) / WY_ROWS) * ((bl[i].m + 63) / 64);
    if (g < cnt) {
      b = i;
      t = g;
      return true;
    }
    g -= cnt;
  }
  return false;
}
struct WyPart {
  int n, m, j0, nbp, R, C, c, ct;
};
__device__ __forceinline__ WyPart ht_wy_part(const HtBlock& B, int step, int t) {
  WyPart q;
  q.n = B.n;
  q.m = B.m;
  const int P = (q.n - 1 + HT_NB - 1) / HT_NB;
  q.j0 = (P - 1 - step) * HT_NB;
  q.nbp = min(HT_NB, q.n - 1 - q.j0);
  q.R = q.n - q.j0 - 1;
  q.C = (q.R + WY_ROWS - 1) / WY_ROWS;
  const int CT = (q.m + 63) / 64;
  q.c = t / CT;
  q.ct = t % CT;
  return q;
}
__global__ void __launch_bounds__(256) k_ht_wy_y(const HtBlock* __restrict__ bl, int nblk, const HtLaunch L, int step,
                                                 const double* __restrict__ Va, const double* __restrict__ Ua,
                                                 double* __restrict__ Ys) {
  int b, t;
  if (!ht_wy_find(bl, nblk, L, step, blockIdx.x, b, t)) return;
  const HtBlock B = bl[b];
  const WyPart q = ht_wy_part(B, step, t);
  const double* V = Va + B.p_off;
  const double* U = Ua + B.a_off;
  const long long ld = B.ld;
  // wave wv: panel columns jb .. jb + 15 of Y, tile columns cb .. cb + 31 (two 16 x 16 blocks), summed
  // over the chunk's rows four at a time on the FP64 matrix cores.  Lane l supplies V(r, j0 + jb + l%16)
  // and Z(r, c + l%16) for row r = r0 + l/16 of a 4-row step, straight from memory (16 lanes read 128
  // consecutive bytes of a row); entry (j, c) of a block lands in lane 16 (j%4) + c, register j/4.
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
  const int jb = (wv & 1) * 16, c0 = q.ct * 64 + (wv >> 1) * 32 + li, c1 = c0 + 16;
  const int row0 = q.j0 + 1 + q.c * WY_ROWS, rend = min(row0 + WY_ROWS, q.n);
  const bool jok = jb + li < q.nbp, ok0 = c0 < q.m, ok1 = c1 < q.m;
  const double* vp = V + q.j0 + jb + li;
  dbl4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  for (int r0 = row0; r0 < rend; r0 += 32) {
    double a[8], b0[8], b1[8];
#pragma unroll
    for (int s = 0; s < 8; s++) {
      const int r = r0 + 4 * s + lk;
      const bool rok = r < rend;
      a[s] = rok && jok ? vp[r * ld] : 0.0;
      b0[s] = rok && ok0 ? U[(long long)r * q.n + c0] : 0.0;
      b1[s] = rok && ok1 ? U[(long long)r * q.n + c1] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < 8; s++) {
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b0[s], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b1[s], acc1, 0, 0, 0);
    }
  }
  double* P = Ys + B.s_off + (long long)q.c * HT_NB * q.m;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int j = jb + 4 * r + lk;
    if (j < q.nbp) {
      if (ok0) P[(long long)j * q.m + c0] = acc0[r];
      if (ok1) P[(long long)j * q.m + c1] = acc1[r];
    }
  }
}
__global__ void __launch_bounds__(256) k_ht_wy_z(const HtBlock* __restrict__ bl, int nblk, const HtLaunch L, int step,
                                                 const double* __restrict__ VTa, const double* __restrict__ Ys,
                                                 double* __restrict__ Ua) {
  int b, t;
  if (!ht_wy_find(bl, nblk, L, step, blockIdx.x, b, t)) return;
  const HtBlock B = bl[b];
  const WyPart q = ht_wy_part(B, step, t);
  const double* VT = VTa + B.p_off;
  double* U = Ua + B.a_off;
  const long long ld = B.ld;
  __shared__ double sY[HT_NB][64];
  const int tid = threadIdx.x;
  // Y's tile: the chunks' partials summed in chunk order (thread: 8 entries, one load of each per chunk)
  {
    const double* P = Ys + B.s_off;
    double y[8];
#pragma unroll
    for (int i = 0; i < 8; i++) y[i] = 0.0;
    for (int c = 0; c < q.C; c++) {
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int e = tid + 256 * i, j = e >> 6, col = q.ct * 64 + (e & 63);
        if (j < q.nbp && col < q.m) y[i] += P[((long long)c * HT_NB + j) * q.m + col];
      }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int e = tid + 256 * i;
      sY[e >> 6][e & 63] = y[i];
    }
  }
  // wave wv: the chunk's rows 64 wv .. 64 wv + 63 (four 16-row blocks) x the tile's 64 columns (four
  // blocks).  The accumulators start as Z and take -(V_p T_p)(r, j0 + j) Y(j, c) on the matrix cores,
  // four panel columns a step: lane l supplies -(V_p T_p)(r + l%16, j0 + j + l/16) and Y(j + l/16, c + l%16)
  const int wv = tid >> 6, l = tid & 63, li = l & 15, lk = l >> 4;
  const int rw = q.j0 + 1 + q.c * WY_ROWS + 64 * wv;
  dbl4 acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = rw + 16 * bi + 4 * r + lk, col = q.ct * 64 + 16 * bj + li;
        acc[bi][bj][r] = row < q.n && col < q.m ? U[(long long)row * q.n + col] : 0.0;
      }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < HT_NB / 4; s++) {
    const int j = 4 * s + lk;
    double a[4], y[4];
#pragma unroll
    for (int bi = 0; bi < 4; bi++) {
      const int row = rw + 16 * bi + li;
      a[bi] = row < q.n && j < q.nbp ? -VT[row * ld + q.j0 + j] : 0.0;
    }
#pragma unroll
    for (int bj = 0; bj < 4; bj++) y[bj] = sY[j][16 * bj + li];
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++) acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], y[bj], acc[bi][bj], 0, 0, 0);
  }
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = rw + 16 * bi + 4 * r + lk, col = q.ct * 64 + 16 * bj + li;
        if (row < q.n && col < q.m) U[(long long)row * q.n + col] = acc[bi][bj][r];
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
                   hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
  const int g = ht_grid(bl, k, 0);
  if (g && t0) {
    // the events take the kernel's own start and end, not the enqueueing around it
    thread_counts().launches++;
    hipExtLaunchKernelGGL(k_ht_symv, dim3(g), dim3(256), 0, s, t0, t1, 0, d_bl, (int)bl.size(),
                          ht_launch_table(bl, k, 0), k, A, V, W);
  } else if (g) {
    NRGK_LAUNCH(k_ht_symv, dim3(g), dim3(256), 0, s, d_bl, (int)bl.size(), ht_launch_table(bl, k, 0), k, A, V, W);
  }
  NRGK_HIP(hipGetLastError());
  return g;
}
int launch_ht_update(const HtBlock* d_bl, const std::vector<HtBlock>& bl, int k, const double* A, double* V, double* W,
                     hipStream_t s) {
  const int g = ht_grid(bl, k, 1);
  if (g) NRGK_LAUNCH(k_ht_update, dim3(g), dim3(320), 0, s, d_bl, (int)bl.size(), ht_launch_table(bl, k, 1), k, A, V, W);
  NRGK_HIP(hipGetLastError());
  return g;
}
int launch_ht_syr2k(const HtBlock* d_bl, const std::vector<HtBlock>& bl, int k, double* A, const double* V, double* W,
                    hipStream_t s) {
  const int g = ht_grid(bl, k, 2);
  if (g) NRGK_LAUNCH(k_ht_syr2k, dim3(g), dim3(256), 0, s, d_bl, (int)bl.size(), ht_launch_table(bl, k, 2), k, A, V, W);
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
                     double* S, double ctol_rel, int* list, int* count, hipStream_t s) {
  int g = 0, nmax = 1;
  for (const auto& b : bl) {
    g += (b.m + HT_CHUNK - 1) / HT_CHUNK;
    nmax = std::max(nmax, b.n);
  }
  if (!g) return;
  const int nb = (int)bl.size();
  const size_t tl = (size_t)nmax * 2 * sizeof(double);  // d and e of the widest block
  NRGK_HIP(hipMemsetAsync(count, 0, sizeof(int), s));
  NRGK_LAUNCH(k_ht_stein_clusters, dim3(nb), dim3(256), 0, s, d_bl, W, lam, ctol_rel, list, count);
  for (int it = 0; it <= HT_ITERS; it++) {
    if (it < HT_ITERS)
      NRGK_LAUNCH(k_ht_stein_step, dim3(g), dim3(64), tl, s, d_bl, nb, W, lam, U, S, ctol_rel, it == 0 ? 1 : 0);
    // a fixed grid that walks the cluster list (its length is read on the device)
    NRGK_LAUNCH(k_ht_stein_mgs, dim3(1024), dim3(256), 0, s, d_bl, (const int*)list, (const int*)count, U);
  }
  NRGK_HIP(hipGetLastError());
}
int ht_stein_slots(int m) { return (m + HT_CHUNK - 1) / HT_CHUNK * 64; }
// the back-transformation step `step` (panel P-1-step of every block that has it): partials, then Z
int launch_ht_wy(const HtBlock* d_bl, const std::vector<HtBlock>& bl, int step, const double* V, const double* VT,
                 double* U, double* Ys, hipStream_t s) {
  HtLaunch L{};
  int g = 0;
  L.nb = (int)bl.size() <= HT_LMAX ? 0 : -1;
  for (const auto& b : bl) {
    const int P = (b.n - 1 + HT_NB - 1) / HT_NB;
    if (step < P) {
      const int R = b.n - (P - 1 - step) * HT_NB - 1;
      g += ((R + WY_ROWS - 1) / WY_ROWS) * ((b.m + 63) / 64);
    }
    if (L.nb >= 0) L.end[L.nb++] = g;
  }
  if (!g) return 0;
  const int nb = (int)bl.size();
  NRGK_LAUNCH(k_ht_wy_y, dim3(g), dim3(256), 0, s, d_bl, nb, L, step, V, (const double*)U, Ys);
  NRGK_LAUNCH(k_ht_wy_z, dim3(g), dim3(256), 0, s, d_bl, nb, L, step, VT, (const double*)Ys, U);
  NRGK_HIP(hipGetLastError());
  return 2;
}
void launch_ht_larft(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* V, const double* W, double* A,
                     hipStream_t s) {
  int g = 0;
  for (const auto& b : bl) g += (b.n - 1 + HT_NB - 1) / HT_NB;
  if (g) NRGK_LAUNCH(k_ht_larft, dim3(g), dim3(256), 0, s, d_bl, (int)bl.size(), V, W, A);
  NRGK_HIP(hipGetLastError());
}
// dynamic LDS of the bisection and inverse-iteration kernels (d and e^2, or d and e, of the widest
// block); their static LDS (below 8 KB) comes on top of it
constexpr int HT_STATIC_LDS = 8192;
int ht_max_n(int lds_bytes) { return (lds_bytes - HT_STATIC_LDS) / (int)(2 * sizeof(double)); }
void set_ht_lds_limits(int bytes) {
  const int dyn = bytes - HT_STATIC_LDS;
  NRGK_HIP(hipFuncSetAttribute((const void*)k_ht_stebz, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
  NRGK_HIP(hipFuncSetAttribute((const void*)k_ht_stein_step, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
}

}  // namespace nrgk
