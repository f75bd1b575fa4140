// K2/K6: batched symmetric eigensolver replacing per-block LAPACKE_dsyevd
// (nrgcore/include/utils/qmatrix.hpp:375-404; called serially per sector at nrgcore.hpp:85-89).
//
// Stage 1 (FP32): block one-sided (Hestenes) Jacobi on W = H + sigma I (positive definite, so
//   W converges to V diag(lambda + sigma) and no separate V is kept).  Block columns of width bw
//   are paired by a round-robin tournament; one workgroup owns one pair: it stages the n x 2bw
//   panel in LDS, runs inner Jacobi sweeps over the panel's columns (16 waves = 16 column pairs in
//   flight, dot products accumulated in FP64) and writes the panel back.  All blocks of an
//   iteration advance one tournament round per launch.
// Stage 2 (FP64): Ogita-Aishima refinement X <- X + X E from S = X^T H X and R = X^T X, with
//   clusters of close eigenvalues resolved by a Rayleigh-Ritz step in FP64 (one-sided Jacobi on the
//   cluster's k x k matrix), then eigenvalues sorted ascending with their columns, as dsyevd
//   returns them (eigenvectors in the columns of U, qmatrix.hpp:390-391).
#include "common.hpp"
#include "kernels.hpp"

namespace nrgk {

constexpr int JAC_THREADS = 1024;
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// tournament_pair without integer division: j < N/2 and round < N-1 keep both sums below 2(N-1)
__device__ __forceinline__ void tournament_pair_fast(int N, int round, int j, int& p, int& q) {
  const int m = N - 1;
  if (j == 0) {
    p = 0;
    q = 1 + round;
  } else {
    int x = j + round;
    x -= x >= m ? m : 0;
    int y = m - j + round;
    y -= y >= m ? m : 0;
    p = 1 + x;
    q = 1 + y;
  }
}

__device__ __forceinline__ void tournament_pair(int N, int round, int j, int& p, int& q) {
  // N players (even); round in [0, N-1); pair j in [0, N/2)
  if (j == 0) {
    p = 0;
    q = 1 + (round % (N - 1));
  } else {
    p = 1 + (j + round) % (N - 1);
    q = 1 + (N - 1 - j + round) % (N - 1);
  }
}

// One tournament round for the tasks of the blocks still rotating.  flags[2b] is set when block b
// rotates during the sweep, flags[2b + 1] when some pair of it starts a round more than JAC_FAR from
// orthogonal (cosine of the angle between the columns; far2 = JAC_FAR^2).  The host drops a block after
// a sweep that leaves it unrotated, or whose pairs all started within JAC_FAR: Jacobi converges
// quadratically, so that sweep has taken them to about JAC_FAR^2, below the rotation tolerance, and
// a further sweep would only find that out.
//
// A task owns the pair (p, q) of block columns of W (column-major, length n).  Its local columns
// are p's at 0..wp-1 and q's at bw..bw+wq-1 (bw <= 16; the columns between exist only beside the
// last, partial block column and are zero).  Instead of rotating the columns pair by pair, it forms
// their Gram matrix G = P^T P in FP64 on the matrix cores (P read once, staged through LDS in row
// chunks), runs the cyclic Jacobi rotations on G (two-sided, the same rotations a one-sided sweep of
// P would make) while accumulating V, and finally writes P <- P V in one pass, also on the matrix
// cores.  The rounding of P to FP32 then happens once per round, not per rotation.
constexpr int JAC_BW = 16;
constexpr int JAC_W = 2 * JAC_BW;
constexpr int JAC_CHUNK = 768;         // rows per LDS chunk of the Gram pass
constexpr int JAC_CS = JAC_CHUNK + 4;  // column stride of the chunk: 16 columns x 4 rows fall in distinct banks
// G and V carry a zero row and column JAC_DUMMY, read in place of a missing member of a rotation
// pair, and a row and column JAC_SINK that take its writes, so that a rotation step has no branches
constexpr int JAC_DUMMY = JAC_W, JAC_SINK = JAC_W + 1;
constexpr int GP = JAC_W + 3;          // row stride of G and V (rows and columns 0..JAC_SINK)
constexpr int GSZ = (JAC_W + 2) * GP;

// sum of the four partial 16 x 16 blocks at P, P + 256, P + 512, P + 768
__device__ __forceinline__ double sum4(const double* P, int t) { return (P[t] + P[256 + t]) + (P[512 + t] + P[768 + t]); }

__global__ void __launch_bounds__(JAC_THREADS) k_jacobi_round(const JacTask* __restrict__ tasks, float* __restrict__ W,
                                                              int round, int* __restrict__ flags,
                                                              double* __restrict__ dgram, int* __restrict__ dgv,
                                                              float tol, float ang_tol, int max_inner, double far2) {
  extern __shared__ float4 chunk4[];  // [JAC_W * JAC_CS] floats, column-major, 16-byte aligned
  float* chunk = reinterpret_cast<float*>(chunk4);
  __shared__ double G[GSZ], G2[GSZ], Vd[GSZ];
  __shared__ double Gpart[3 * 1024];  // partial Gram blocks of the matrix-core waves
  __shared__ int s_slots[32 * 16];
  __shared__ float Vf[JAC_W * GP];
  __shared__ int s_any;
  const JacTask task = tasks[blockIdx.x];
  const int b = task.b, j = task.j;
  const int N = task.nbc;
  if (round >= N - 1) return;
  int p, q;
  tournament_pair(N, round, j, p, q);
  const int n = task.n, bw = task.bw;
  const int tid = threadIdx.x;
  const int c0p = p * bw, c0q = q * bw;
  const int wp = c0p < n ? min(bw, n - c0p) : 0;
  const int wq = c0q < n ? min(bw, n - c0q) : 0;
  if (wp + wq == 0) return;
  // beside the empty block column that pads the tournament to an even count there is nothing to
  // pair after round 0 (which rotates the pairs inside each block column)
  if (round > 0 && (wp == 0 || wq == 0)) return;
  const int wl = bw + wq;  // local columns in use: [0, wp) and [bw, wl)
  float* Wb = W + task.off;
  auto valid = [&](int i) { return i < wp || (i >= bw && i < wl); };
  auto gcol = [&](int i) { return i < bw ? c0p + i : c0q + (i - bw); };
  const bool resident = n <= JAC_CHUNK;
  // Gram blocks of the two block columns (16 x 16 slots, FP64), kept from their last round: when both
  // are valid, only the cross block is formed from P.  They are the Gram blocks of the exact rotated
  // columns, which the FP32 panel holds to rounding.  A block column keeps its block current in every
  // round that touches it, so a stored block is never stale.  Their validity and the blocks are read
  // together (both slots exist whatever their state).
  double* dgp = dgram + ((long long)task.bc_off + p) * 256;
  double* dgq = dgram + ((long long)task.bc_off + q) * 256;
  const int vp = dgv[task.bc_off + p], vq = dgv[task.bc_off + q];
  double pre_p = 0.0, pre_q = 0.0;
  if (tid < 256) {
    pre_p = dgp[tid];
    pre_q = dgq[tid];
  }
  const bool keepP = wp > 0, keepQ = wq > 0;
  const bool fast = resident && keepP && keepQ && vp && vq;
  // ---- step schedule of the Jacobi rotations.  A step rotates up to 16 disjoint column pairs
  // ("slots").  Each pair of columns is rotated once per outer sweep: the pairs inside p and inside q
  // in round 0 (a circle tournament of bwp = bw rounded up to even players on each block at once,
  // bwp - 1 steps, slots 0..bwp/2-1 on p and bwp/2..bwp-1 on q), the bw x bw cross pairs in every
  // round (step k, slot i: p-column i with q-column (i + k) mod bw).  A column without a partner in
  // a step sits in a slot with -1.
  const int bwp = bw + (bw & 1);
  const int n_intra = round == 0 ? bwp - 1 : 0;
  const int steps_per = n_intra + bw;
  const int nsteps = max_inner * steps_per;
  auto slot = [&](int st, int sl, int& u, int& v) {
    if (st >= n_intra) {  // cross step (uniform branch)
      int jj = sl + st - n_intra;
      jj -= jj >= bw ? bw : 0;
      u = sl < wp ? sl : -1;
      v = sl < bw && jj < wq ? bw + jj : -1;
      return;
    }
    u = v = -1;
    const int half = bwp >> 1, M = bwp - 1;
    const bool sq = sl >= half;
    const int m = sq ? sl - half : sl;
    if (m >= half) return;
    int a = 0, c = 1 + st;
    if (m > 0) {
      int x = st + m;
      x -= x >= M ? M : 0;
      int y = st - m;
      y += y < 0 ? M : 0;
      a = 1 + x;
      c = 1 + y;
    }
    if (a > c) {
      const int t = a;
      a = c;
      c = t;
    }
    const int wb = sq ? wq : wp, base = sq ? bw : 0;
    u = a < wb ? base + a : -1;
    v = c < wb ? base + c : -1;
  };
  // set-up that needs no data, done while the panel loads are in flight: the slot table (one period,
  // steps_per x 16: read and write indices of the pair packed ru | rv << 8 | wu << 16 | wv << 24, a
  // missing member reading JAC_DUMMY and writing JAC_SINK), V = I, and the dummy row and column of
  // both G buffers set to zero
  auto setup = [&]() {
    if (tid == 0) s_any = 0;
    for (int e = tid; e < steps_per * 16; e += JAC_THREADS) {
      int u, v;
      slot(e >> 4, e & 15, u, v);
      const int ru = u >= 0 ? u : JAC_DUMMY, rv = v >= 0 ? v : JAC_DUMMY;
      const int wu = u >= 0 ? u : JAC_SINK, wv2 = v >= 0 ? v : JAC_SINK;
      s_slots[e] = ru | (rv << 8) | (wu << 16) | (wv2 << 24);
    }
    for (int e = tid; e < (JAC_W + 2) * (JAC_W + 2); e += JAC_THREADS) {
      const int x = e / (JAC_W + 2), y = e - x * (JAC_W + 2);
      Vd[x * GP + y] = x == y && x < JAC_W ? 1.0 : 0.0;
      if (x == JAC_DUMMY || y == JAC_DUMMY) {
        G[x * GP + y] = 0.0;
        G2[x * GP + y] = 0.0;
      }
    }
  };
  // ---- Gram matrix on the FP64 matrix cores (16 x 16 x 4 steps).  Job of wave v: the p x q block
  // (fast: waves 0-11, each taking every twelfth group of 4 rows), or the p x p, p x q and q x q
  // blocks (waves 0-3, 4-7, 8-11, every fourth group); two accumulators in flight.  Lane l supplies column
  // l%16 of its left block as A and of its right block as B, at row 4 kb + l/16; entry (i, j) of the
  // product lands in lane 16 (i%4) + j, register i/4.  The partial blocks meet in Gpart.
  const int wv = tid >> 6, lane = tid & 63, lx = lane & 15;
  const int job = wv < 12 ? (fast ? 1 : wv >> 2) : -1;
  const int kq = fast ? wv : wv & 3, kstride = fast ? 12 : 4;
  const float* ca = chunk + (job == 2 ? bw + lx : lx) * JAC_CS + (lane >> 4);
  const float* cb = chunk + (job == 0 ? lx : bw + lx) * JAC_CS + (lane >> 4);
  // a 16-wide tile of p reaches into q when bw < 16: those lanes supply zeros
  const bool am = job == 2 || lx < bw, bm = job != 0 || lx < bw;
  dbl4 m0 = {0.0, 0.0, 0.0, 0.0}, m1 = {0.0, 0.0, 0.0, 0.0};
  for (int r0 = 0; r0 < n; r0 += JAC_CHUNK) {
    const int rows = min(JAC_CHUNK, n - r0);
    const int rows4 = (rows + 3) & ~3;
    // thread = (column tid/32, rows tid%32 + 32k): one column pointer per thread, 128-byte runs, all
    // of a chunk's loads in flight before their LDS stores; rows are filled up to a multiple of 4
    {
      const int i = tid >> 5, l = tid & 31;
      const bool ok = valid(i);
      const float* src = Wb + (long long)gcol(ok ? i : 0) * n + r0 + l;
      float* dst = chunk + i * JAC_CS + l;
      const int lim = ok ? rows - l : 0, lim4 = rows4 - l;
      float t[JAC_CHUNK / 32];
#pragma unroll
      for (int k = 0; k < JAC_CHUNK / 32; k++) t[k] = 32 * k < lim ? src[32 * k] : 0.0f;
#pragma unroll
      for (int k = 0; k < JAC_CHUNK / 32; k++)
        if (32 * k < lim4) dst[32 * k] = t[k];
    }
    if (r0 == 0) setup();
    if (fast && tid < 256) {  // (resident: this is the only chunk); entries outside a partial block
      const int a = tid >> 4, c = tid & 15;  // column's width are zero
      if (a < bw && c < bw) {
        G[a * GP + c] = a < wp && c < wp ? pre_p : 0.0;
        G[(bw + a) * GP + bw + c] = a < wq && c < wq ? pre_q : 0.0;
      }
    }
    __syncthreads();
    if (job >= 0) {
      const int nkb = rows4 >> 2;
      int kb = kq;
      // four groups per iteration: all eight operands are read before the products
      for (; kb + 3 * kstride < nkb; kb += 4 * kstride) {
        double xa[4], xb[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          xa[u] = am ? (double)ca[4 * (kb + u * kstride)] : 0.0;
          xb[u] = bm ? (double)cb[4 * (kb + u * kstride)] : 0.0;
        }
        m0 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[0], xb[0], m0, 0, 0, 0);
        m1 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[1], xb[1], m1, 0, 0, 0);
        m0 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[2], xb[2], m0, 0, 0, 0);
        m1 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[3], xb[3], m1, 0, 0, 0);
      }
      for (; kb < nkb; kb += kstride)
        m0 = __builtin_amdgcn_mfma_f64_16x16x4f64(am ? (double)ca[4 * kb] : 0.0, bm ? (double)cb[4 * kb] : 0.0, m0,
                                                  0, 0, 0);
    }
    if (!resident) __syncthreads();  // the next chunk overwrites this one
  }
  if (job >= 0) {
    double* part = Gpart + (fast ? 0 : job * 1024) + kq * 256 + (lane >> 4) * 16 + lx;
#pragma unroll
    for (int r = 0; r < 4; r++) part[64 * r] = m0[r] + m1[r];
  }
  __syncthreads();
  if (tid < 256) {
    const int a = tid >> 4, c = tid & 15;
    if (a < bw && c < bw) {
      const double gx =
          fast ? (sum4(Gpart, tid) + sum4(Gpart + 1024, tid)) + sum4(Gpart + 2048, tid) : sum4(Gpart + 1024, tid);
      G[a * GP + bw + c] = gx;
      G[(bw + c) * GP + a] = gx;
      if (!fast) {
        G[a * GP + c] = sum4(Gpart, tid);
        G[(bw + a) * GP + bw + c] = sum4(Gpart + 2048, tid);
      }
    }
  }
  __syncthreads();
  if (!fast) {  // keep the freshly formed diagonal blocks
    if (tid < 256) {
      const int a = tid >> 4, c = tid & 15;
      if (a < bw && c < bw) {
        if (keepP) dgp[tid] = G[a * GP + c];
        if (keepQ) dgq[tid] = G[(bw + a) * GP + bw + c];
      }
    }
    if (tid == 0) {
      if (keepP) dgv[task.bc_off + p] = 1;
      if (keepQ) dgv[task.bc_off + q] = 1;
    }
  }
  {
    // a task whose columns are already orthogonal to the tolerance has nothing to rotate (the zero
    // columns between p and q have zero entries)
    const int x = tid >> 5, y = tid & 31;
    bool need = false, far = false;
    if (x < y && y < wl) {
      const double ga = G[x * GP + y], ab = G[x * GP + x] * G[y * GP + y];
      need = ga * ga > ((double)tol * (double)tol) * ab;
      far = ga * ga > far2 * ab;
    }
    // one barrier for both tests; flags[2b + 1]: some pair of the block was at least JAC_FAR from
    // orthogonal during the sweep
    const unsigned long long bneed = __ballot(need), bfar = __ballot(far);
    if ((tid & 63) == 0 && (bneed | bfar)) atomicOr(&s_any, (bneed ? 1 : 0) | (bfar ? 2 : 0));
    __syncthreads();
    const int any = s_any;
    if ((any & 2) && tid == 0) flags[2 * b + 1] = 1;
    if (!(any & 1)) return;
  }
  // ---- cyclic Jacobi on G.  Column x of slot (u, v), u < v, moves as P'(:,u) = c P(:,u) - s P(:,v),
  // P'(:,v) = s P(:,u) + c P(:,v), so G' = J^T G J and V' = V J.  Every thread that needs a slot's
  // rotation computes it from the same three entries of the current G (so all copies agree exactly).
  // One barrier per step.
  const double tol2 = (double)tol * (double)tol;
  // rotation of slot (u, v) (read indices) from the current G: identity for a lone column (its
  // partner is the zero dummy, ga = 0) or a pair below tolerance.  Branch-free: a zero or tiny ga
  // gives inf/NaN on the discarded side of the final selection only.
  auto rot = [&](const double* Gc, int u, int v, double& cs, double& sn) {
    const double al = Gc[u * GP + u], be = Gc[v * GP + v], ga = Gc[u * GP + v];
    // the angle to FP32 accuracy; cos and sin then normalised in FP64 (cs^2 + sn^2 = 1 to rounding,
    // one Newton step on the FP32 reciprocal square root), so the accumulated V stays orthogonal
    const float zeta = (float)(be - al) * __builtin_amdgcn_rcpf((float)(2.0 * ga));
    const float tf = __builtin_copysignf(1.0f, zeta) *
                     __builtin_amdgcn_rcpf(fabsf(zeta) + __builtin_amdgcn_sqrtf(1.0f + zeta * zeta));
    const double t = (double)tf, xx = 1.0 + t * t;
    double z = (double)__builtin_amdgcn_rsqf((float)xx);
    z = z * (1.5 - 0.5 * xx * z * z);
    // a rotation below the angle tolerance changes the vectors by less than FP32 can hold
    const bool go = (ga != 0.0) & (ga * ga > tol2 * (al * be)) & (fabsf(tf) > ang_tol);
    cs = go ? z : 1.0;
    sn = go ? z * t : 0.0;
  };
  // thread (a, c) < 256 rotates the entries (rows of slot a) x (columns of slot c) of G into the
  // other buffer, and the same entries of V in place (V' = V J needs only slot c's rotation); the
  // slots partition the columns, so every entry is written once.  Lane l computes the rotation of
  // slot l%16, so slot a's comes from lane a of the same wave.  Threads 256 and up only keep the
  // barrier.
  double* Gc = G;
  double* Gn = G2;
  const int sa_i = tid >> 4, sc_i = tid & 15;
  bool rotated = false;
  int ea = 0, ec = 0;
  if (tid < 256) {
    ea = s_slots[sa_i];
    ec = s_slots[sc_i];
  }
  for (int st = 0, row = 0; st < nsteps; st++) {
    row = row + 1 == steps_per ? 0 : row + 1;  // the next step's row of the table
    if (tid < 256) {
      const int rua = ea & 0xff, rva = (ea >> 8) & 0xff, wua = (ea >> 16) & 0xff, wva = ea >> 24;
      const int ruc = ec & 0xff, rvc = (ec >> 8) & 0xff, wuc = (ec >> 16) & 0xff, wvc = ec >> 24;
      ea = s_slots[row * 16 + sa_i];
      ec = s_slots[row * 16 + sc_i];
      double cc, sc;
      rot(Gc, ruc, rvc, cc, sc);
      const double g00 = Gc[rua * GP + ruc], g01 = Gc[rua * GP + rvc], g10 = Gc[rva * GP + ruc], g11 = Gc[rva * GP + rvc];
      const double v00 = Vd[rua * GP + ruc], v01 = Vd[rua * GP + rvc], v10 = Vd[rva * GP + ruc], v11 = Vd[rva * GP + rvc];
      const double ca = __shfl(cc, sa_i, 64), sa = __shfl(sc, sa_i, 64);
      rotated |= sc != 0.0;
      // columns, then rows
      const double r00 = cc * g00 - sc * g01, r01 = sc * g00 + cc * g01;
      const double r10 = cc * g10 - sc * g11, r11 = sc * g10 + cc * g11;
      Gn[wua * GP + wuc] = ca * r00 - sa * r10;
      Gn[wua * GP + wvc] = ca * r01 - sa * r11;
      Gn[wva * GP + wuc] = sa * r00 + ca * r10;
      Gn[wva * GP + wvc] = sa * r01 + ca * r11;
      Vd[wua * GP + wuc] = cc * v00 - sc * v01;
      Vd[wua * GP + wvc] = sc * v00 + cc * v01;
      Vd[wva * GP + wuc] = cc * v10 - sc * v11;
      Vd[wva * GP + wvc] = sc * v10 + cc * v11;
    }
    __syncthreads();
    double* t = Gc;
    Gc = Gn;
    Gn = t;
  }
  const bool any_rot = __syncthreads_or(rotated);
  if (!any_rot) return;  // P unchanged: nothing to write
  const int ex = tid >> 5, ey = tid & 31;
  Vf[ex * GP + ey] = (float)Vd[ex * GP + ey];
  {  // the diagonal blocks of G' = V^T G V for the next round
    const double gv = Gc[ex * GP + ey];
    if (keepP && ex < wp && ey < wp) dgp[ex * JAC_BW + ey] = gv;
    if (keepQ && ex >= bw && ey >= bw && ex < wl && ey < wl) dgq[(ex - bw) * JAC_BW + (ey - bw)] = gv;
  }
  __syncthreads();
  // ---- P <- P V in FP32
  if (resident) {
    // (P V)^T = V^T P^T on the FP32 matrix cores in 16-row tiles: per tile and half h of the new
    // columns, 8 steps of 16 x 16 x 4 over the 32 old columns.  Lane l keeps V(4 ks + l/16, 16 h + l%16)
    // (A) in registers and reads P(row0 + l%16, 4 ks + l/16) (B); entry (i, j) = (P V)(row0 + j,
    // 16 h + i) lands in lane 16 (i/4) + j, register i%4 and goes straight to W (16 consecutive rows
    // of 4 columns per store).  Tile rows past n read the column's padding and are not stored.
    const int li = lane & 15, lk = lane >> 4;
    float va[2][8];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int ks = 0; ks < 8; ks++) va[h][ks] = Vf[(4 * ks + lk) * GP + 16 * h + li];
    // the eight columns this lane stores: 16 h + 4 (l/16) + r (absent ones flagged)
    float* colp[2][4];
    bool colv[2][4];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = 16 * h + 4 * lk + r;
        colv[h][r] = valid(col);
        colp[h][r] = Wb + (long long)gcol(colv[h][r] ? col : 0) * n;
      }
    auto store = [&](const flt4& m, int row, int h) {
      if (row < n)
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (colv[h][r]) colp[h][r][row] = m[r];
    };
    // two tiles per pass (waves take tiles v, v + 16, v + 32, ...): four independent accumulator
    // chains keep the matrix pipeline busy
    const int ntile = (n + 15) >> 4, nw = JAC_THREADS / 64;
    for (int t0 = wv; t0 < ntile; t0 += 2 * nw) {
      const int t1 = t0 + nw;
      const int row0 = 16 * t0 + li, row1 = 16 * t1 + li;
      const bool two = t1 < ntile;
      float pb0[8], pb1[8];
#pragma unroll
      for (int ks = 0; ks < 8; ks++) {
        pb0[ks] = chunk[(4 * ks + lk) * JAC_CS + row0];
        pb1[ks] = two ? chunk[(4 * ks + lk) * JAC_CS + row1] : 0.0f;
      }
      flt4 m00 = {0.0f, 0.0f, 0.0f, 0.0f}, m01 = m00, m10 = m00, m11 = m00;
#pragma unroll
      for (int ks = 0; ks < 8; ks++) {
        m00 = __builtin_amdgcn_mfma_f32_16x16x4f32(va[0][ks], pb0[ks], m00, 0, 0, 0);
        m01 = __builtin_amdgcn_mfma_f32_16x16x4f32(va[1][ks], pb0[ks], m01, 0, 0, 0);
        m10 = __builtin_amdgcn_mfma_f32_16x16x4f32(va[0][ks], pb1[ks], m10, 0, 0, 0);
        m11 = __builtin_amdgcn_mfma_f32_16x16x4f32(va[1][ks], pb1[ks], m11, 0, 0, 0);
      }
      store(m00, row0, 0);
      store(m01, row0, 1);
      if (two) {
        store(m10, row1, 0);
        store(m11, row1, 1);
      }
    }
    if (tid == 0) flags[2 * b] = 1;
    return;
  }
  // larger panels: one row per thread from global memory (every value of the row is read before it
  // is overwritten)
  for (int r = tid; r < n; r += JAC_THREADS) {
    float x[JAC_W];
#pragma unroll
    for (int i = 0; i < JAC_W; i++) x[i] = valid(i) ? Wb[(long long)gcol(i) * n + r] : 0.0f;
    for (int c = 0; c < JAC_W; c++) {
      if (!valid(c)) continue;
      float s = 0.0f;
#pragma unroll
      for (int i = 0; i < JAC_W; i++) s = fmaf(x[i], Vf[i * GP + c], s);
      Wb[(long long)gcol(c) * n + r] = s;
    }
  }
  if (tid == 0) flags[2 * b] = 1;
}

// Column norms of the converged W give lambda + sigma; columns are normalised, sorted by lambda
// and written as FP64 X (row-major, eigenvectors in columns).
// Column norms of W, 32 columns of one block per workgroup (one wave per column at a time)
__global__ void __launch_bounds__(256) k_col_norms(const EigBlock* __restrict__ blocks, const int4* __restrict__ groups,
                                                   const float* __restrict__ W, double* __restrict__ nrm_g) {
  const int4 g = groups[blockIdx.x];
  const EigBlock blk = blocks[g.x];
  const int n = blk.n, c1 = min(g.z + 32, n);
  const float* Wb = W + blk.off;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int c = g.z + wave; c < c1; c += 4) {
    double s = 0.0;
    for (int r = lane; r < n; r += 64) {
      const double v = Wb[(long long)c * n + r];
      s += v * v;
    }
    s = wave_sum(s);
    if (lane == 0) nrm_g[blk.e_off + c] = sqrt(s);
  }
}

__global__ void __launch_bounds__(1024) k_normalize_sort(const EigBlock* __restrict__ blocks, const double* __restrict__ nrm_g,
                                                         const float* __restrict__ sigma, double* __restrict__ lam,
                                                         int* __restrict__ perm, double* __restrict__ nrm_s) {
  extern __shared__ double sh[];
  const EigBlock blk = blocks[blockIdx.x];
  const int n = blk.n;
  if (n == 0) return;
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  double* key = sh;                 // [np2]
  int* idx = (int*)(sh + np2);      // [np2]
  double* nrm = sh + np2 + (np2 + 1) / 2 + 1;  // [n]
  for (int c = threadIdx.x; c < n; c += blockDim.x) nrm[c] = nrm_g[blk.e_off + c];
  __syncthreads();
  const double sg = sigma[blockIdx.x];
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    key[i] = i < n ? nrm[i] - sg : 1e300;
    idx[i] = i;
  }
  __syncthreads();
  // bitonic sort ascending
  for (int k = 2; k <= np2; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = threadIdx.x; i < np2; i += blockDim.x) {
        const int l = i ^ jj;
        if (l > i) {
          const bool up = ((i & k) == 0);
          const double a = key[i], bb = key[l];
          if ((a > bb) == up) {
            key[i] = bb;
            key[l] = a;
            const int t = idx[i];
            idx[i] = idx[l];
            idx[l] = t;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    lam[blk.e_off + c] = key[c];
    perm[blk.e_off + c] = idx[c];
    nrm_s[blk.e_off + c] = nrm[idx[c]];
  }
}

// X(r, c) = W(:, perm c)(r) / |W(:, perm c)| in FP64 (row-major X, eigenvectors in columns), one
// 32 x 32 tile of one block per workgroup: columns are read along W's contiguous direction into LDS
// and the tile is written along X's rows.
__global__ void __launch_bounds__(256) k_gather_x(const EigBlock* __restrict__ blocks, const int4* __restrict__ tiles,
                                                  const float* __restrict__ W, const int* __restrict__ perm,
                                                  const double* __restrict__ nrm_s, double* __restrict__ X) {
  __shared__ float T[32][33];
  const int4 t = tiles[blockIdx.x];
  const EigBlock blk = blocks[t.x];
  const int n = blk.n, r0 = 32 * t.y, c0 = 32 * t.z;
  const float* Wb = W + blk.off;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = threadIdx.x + 256 * k, c = e >> 5, rr = e & 31;
    if (c0 + c < n && r0 + rr < n) T[c][rr] = Wb[(long long)perm[blk.e_off + c0 + c] * n + r0 + rr];
  }
  __syncthreads();
  double* Xb = X + blk.off;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = threadIdx.x + 256 * k, rr = e >> 5, c = e & 31;
    if (c0 + c < n && r0 + rr < n) Xb[(long long)(r0 + rr) * n + c0 + c] = (double)T[c][rr] / nrm_s[blk.e_off + c0 + c];
  }
}

// ---------------------------------------------------------------- Ogita-Aishima update
// Inputs S = X^T H X, Rp = X^T X.  Output E (X <- X + X E) and lam.
//   lt_i = S_ii / Rp_ii; E_ii = (1 - Rp_ii)/2;  E_ij = (S_ij - lt_j Rp_ij)/(lt_j - lt_i)
// for eigenvalues farther apart than delta = max(32 max_{i!=j}|rho_ij| + 64 eps max|lt|,
// 1e5 sqrt(n) eps max|lt|), rho_ij = S_ij - (lt_i + lt_j)/2 Rp_ij.
// Clusters (runs of sorted lt with gaps <= delta) get a Rayleigh-Ritz step: M = 1.5 I - 0.5 Rp_JJ,
// (M S_JJ M) Q = Q mu, E_JJ = M Q - I and the out-of-cluster rows of E(:,J) are mapped through M Q.
constexpr int RR_MAX = 64;

__device__ void rr_jacobi(double* Wm, int k, double* mu, int* perm, double* nrm) {
  // one-sided Jacobi on the SPD k x k matrix Wm (column-major, in LDS); on return the columns are
  // orthogonal: W = Q diag(mu + shift)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __shared__ int s_rot;
  const int kp = (k + 1) & ~1;
  for (int sweep = 0; sweep < 30; sweep++) {
    if (threadIdx.x == 0) s_rot = 0;
    __syncthreads();
    for (int step = 0; step < kp - 1; step++) {
      for (int pj = wave; pj < kp / 2; pj += blockDim.x / 64) {
        int a, c;
        tournament_pair(kp, step, pj, a, c);
        if (a >= k || c >= k) continue;
        double* x = Wm + a * k;
        double* y = Wm + c * k;
        double al = 0, be = 0, ga = 0;
        for (int r = lane; r < k; r += 64) {
          al += x[r] * x[r];
          be += y[r] * y[r];
          ga += x[r] * y[r];
        }
        al = wave_sum(al);
        be = wave_sum(be);
        ga = wave_sum(ga);
        if (ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
          const double zeta = (be - al) / (2.0 * ga);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
          for (int r = lane; r < k; r += 64) {
            const double xv = x[r], yv = y[r];
            x[r] = cs * xv - sn * yv;
            y[r] = sn * xv + cs * yv;
          }
          if (lane == 0) s_rot = 1;
        }
      }
      __syncthreads();
    }
    const int any = s_rot;
    __syncthreads();
    if (!any) break;
  }
  for (int c = wave; c < k; c += blockDim.x / 64) {
    double s = 0;
    for (int r = lane; r < k; r += 64) s += Wm[c * k + r] * Wm[c * k + r];
    s = wave_sum(s);
    if (lane == 0) nrm[c] = sqrt(s);
  }
  __syncthreads();
  // ascending order of the norms by rank (ties by column index)
  for (int c = threadIdx.x; c < k; c += blockDim.x) {
    const double v = nrm[c];
    int r = 0;
    for (int c2 = 0; c2 < k; c2++) {
      const double u = nrm[c2];
      r += (u < v) || (u == v && c2 < c);
    }
    perm[r] = c;
    mu[r] = v;
  }
  __syncthreads();
}

// Block-wide reductions over 1024 threads through red[0..15] (LDS); every thread gets the result.
__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
  __syncthreads();
  return s;
}
__device__ double block_max(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) s = fmax(s, red[w]);
  __syncthreads();
  return s;
}

// Rayleigh-Ritz step for a cluster wider than RR_MAX, with its k x k matrices in global scratch
// (three n x n regions of the block, free while the update runs) instead of LDS.  Same algebra as
// the LDS path: M = 1.5 I - 0.5 Rp_JJ, T = M S_JJ M, one-sided Jacobi on T - cen I + sig I,
// G = M Q, E(i,J) <- E(i,J) G outside the cluster, E_JJ = G - I.  J = members (block indices).
__device__ void rr_wide_cluster(const double* Sb, const double* Rb, double* Eb, int n, int k, const int* J,
                                const int* cid, int c, double* gW, double* gM, double* gT, double* rv, double* lam,
                                double* red, bool map_rows) {
  auto Ss = [&](int i, int j) { return 0.5 * (Sb[(long long)i * n + j] + Sb[(long long)j * n + i]); };
  auto Rs = [&](int i, int j) { return 0.5 * (Rb[(long long)i * n + j] + Rb[(long long)j * n + i]); };
  const int tid = threadIdx.x;
  const long long kk = (long long)k * k;
  for (long long e = tid; e < kk; e += blockDim.x) {
    const int a = (int)(e / k), b = (int)(e % k);
    gM[e] = (a == b ? 1.5 : 0.0) - 0.5 * Rs(J[a], J[b]);
  }
  __syncthreads();
  for (long long e = tid; e < kk; e += blockDim.x) {  // gT = S_JJ M
    const int a = (int)(e / k), b = (int)(e % k);
    double s = 0.0;
    for (int l = 0; l < k; l++) s += Ss(J[a], J[l]) * gM[(long long)l * k + b];
    gT[e] = s;
  }
  __syncthreads();
  for (long long e = tid; e < kk; e += blockDim.x) {  // gW = M (S M)
    const int a = (int)(e / k), b = (int)(e % k);
    double s = 0.0;
    for (int l = 0; l < k; l++) s += gM[(long long)a * k + l] * gT[(long long)l * k + b];
    gW[e] = s;
  }
  __syncthreads();
  double tr = 0.0;
  for (int a = tid; a < k; a += blockDim.x) tr += gW[(long long)a * k + a];
  const double cen = block_sum(tr, red) / k;
  double rmax = 0.0;
  for (int a = tid; a < k; a += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < k; b++) s += fabs(gW[(long long)a * k + b] - (a == b ? cen : 0.0));
    rmax = fmax(rmax, s);
  }
  const double rad = block_max(rmax, red);
  const double sig = rad > 1e-30 * fmax(1.0, fabs(cen)) ? 1.05 * rad : 1.0;
  for (int a = tid; a < k; a += blockDim.x) gW[(long long)a * k + a] = (gW[(long long)a * k + a] - cen) + sig;
  __syncthreads();
  double* mu = rv;
  double* nrm = rv + k;
  int* perm = (int*)(rv + 2 * k);
  rr_jacobi(gW, k, mu, perm, nrm);
  for (long long e = tid; e < kk; e += blockDim.x) {  // gT := G = M Q
    const int a = (int)(e / k), cc = (int)(e % k);
    const int src = perm[cc];
    const double inv = 1.0 / nrm[src];
    double s = 0.0;
    for (int l = 0; l < k; l++) s += gM[(long long)a * k + l] * (gW[(long long)src * k + l] * inv);
    gT[e] = s;
  }
  __syncthreads();
  // old E(:, J) of the rows outside the cluster -> gW (n x k), then E(i, J) = old(i, :) G
  for (long long e = tid; e < (long long)n * k && map_rows; e += blockDim.x) {
    const int i = (int)(e / k), l = (int)(e % k);
    if (cid[i] != c) gW[e] = Eb[(long long)i * n + J[l]];
  }
  __syncthreads();
  for (long long e = tid; e < (long long)n * k && map_rows; e += blockDim.x) {
    const int i = (int)(e / k), cc = (int)(e % k);
    if (cid[i] == c) continue;
    double s = 0.0;
    for (int l = 0; l < k; l++) s += gW[(long long)i * k + l] * gT[(long long)l * k + cc];
    Eb[(long long)i * n + J[cc]] = s;
  }
  for (long long e = tid; e < kk; e += blockDim.x) {
    const int a = (int)(e / k), cc = (int)(e % k);
    Eb[(long long)J[a] * n + J[cc]] = gT[e] - (a == cc ? 1.0 : 0.0);
  }
  for (int cc = tid; cc < k; cc += blockDim.x) lam[J[cc]] = mu[cc] - sig + cen;
  __syncthreads();
}

// Rayleigh-Ritz of one cluster of 2..RR_SMALL members by one wave, with the algebra of the workgroup
// path below: M = 1.5 I - 0.5 Rp_JJ, T = M S_JJ M, one-sided Jacobi on T - cen I + sig I (lane 8a + b
// holds entry (a, b); a column's rows are the lanes b, b + 8, ...), Q and mu ascending, G = M Q,
// E(i, J) <- E(i, J) G for the rows outside the cluster (unless the block takes its last step) and
// E_JJ = G - I.  ws: 3 x 64 doubles of this wave's LDS.
constexpr int RR_SMALL = 8;
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ void rr_small_wave(const double* Sb, const double* Rb, double* Eb, int n, int k, const int* J,
                              const int* cid, int cl, double* lam, double* ws, bool map_rows) {
  const int lane = threadIdx.x & 63, a = lane >> 3, b = lane & 7;
  const bool in = a < k && b < k;
  double* M = ws;
  double* T = ws + 64;
  double* Q = ws + 128;
  const int ja = J[min(a, k - 1)], jb = J[min(b, k - 1)];
  M[lane] = in ? (a == b ? 1.5 : 0.0) - 0.5 * Rb[(long long)ja * n + jb] : 0.0;
  wave_lds_sync();
  double sm = 0.0;  // (S_JJ M)(a, b)
  if (in)
    for (int l = 0; l < k; l++) sm += Sb[(long long)ja * n + J[l]] * M[l * 8 + b];
  T[lane] = sm;
  wave_lds_sync();
  double t = 0.0;  // (M S_JJ M)(a, b)
  if (in)
    for (int l = 0; l < k; l++) t += M[a * 8 + l] * T[l * 8 + b];
  // shift by the cluster centre and make it positive definite (as the workgroup path)
  const double cen = wave_sum(in && a == b ? t : 0.0) / k;
  double ra = in ? fabs(t - (a == b ? cen : 0.0)) : 0.0;
#pragma unroll
  for (int o = 1; o < 8; o <<= 1) ra += __shfl_xor(ra, o, 64);
#pragma unroll
  for (int o = 8; o < 64; o <<= 1) ra = fmax(ra, __shfl_xor(ra, o, 64));
  const double sig = ra > 1e-30 * fmax(1.0, fabs(cen)) ? 1.05 * ra : 1.0;
  double w = in ? (a == b ? (t - cen) + sig : t) : 0.0;
  // one-sided Jacobi on the columns, round-robin over kp players
  const int kp = (k + 1) & ~1;
  auto col_sum = [&](double v) {
#pragma unroll
    for (int o = 8; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
  };
  for (int sweep = 0; sweep < 30; sweep++) {
    bool rot = false;
    for (int step = 0; step < kp - 1; step++) {
      int partner = -1;
      for (int pj = 0; pj < kp / 2; pj++) {
        int x, y;
        tournament_pair(kp, step, pj, x, y);
        if (x == b) partner = y;
        if (y == b) partner = x;
      }
      const bool pair = b < k && partner >= 0 && partner < k;
      const double y = __shfl(w, 8 * a + (pair ? partner : b), 64);
      const double nx = col_sum(w * w), ny = col_sum(y * y), xy = col_sum(w * y);
      // the pair's rotation from (first, second) = (min, max) column.  The sums may differ by an ulp
      // from row to row (a product can be fused into the first add), which at a degenerate pair
      // would change the angle: every lane takes the rotation of row 0 of the first column.
      const bool first = b < partner;
      const double al = first ? nx : ny, be = first ? ny : nx;
      const bool go_l = pair && xy != 0.0 && fabs(xy) > 1e-15 * sqrt(al * be);
      const double zeta = (be - al) / (2.0 * xy);
      const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
      const double cs_l = go_l ? 1.0 / sqrt(1.0 + tt * tt) : 1.0, sn_l = go_l ? cs_l * tt : 0.0;
      const int src = pair ? min(b, partner) : b;
      const double cs = __shfl(cs_l, src, 64), sn = __shfl(sn_l, src, 64);
      const bool go = __shfl(go_l ? 1 : 0, src, 64) != 0;
      if (go) {
        w = first ? cs * w - sn * y : sn * y + cs * w;
        rot = true;
      }
    }
    if (__ballot(rot) == 0) break;
  }
  // columns ascending by norm (ties by index): Q(:, rank) = W(:, b) / |W(:, b)|, mu = |W(:, b)|
  const double nb = __shfl(sqrt(col_sum(w * w)), b, 64);  // row 0's value for the whole column
  int rank = 0;
  for (int c = 0; c < k; c++) {
    const double nc = __shfl(nb, c, 64);
    rank += (nc < nb) || (nc == nb && c < b);
  }
  if (in) Q[a * 8 + rank] = w / nb;
  if (a == 0 && b < k) lam[J[rank]] = nb - sig + cen;
  wave_lds_sync();
  double g = 0.0;  // G = M Q
  if (in)
    for (int l = 0; l < k; l++) g += M[a * 8 + l] * Q[l * 8 + b];
  T[lane] = g;
  wave_lds_sync();
  int jl[RR_SMALL];
#pragma unroll
  for (int l = 0; l < RR_SMALL; l++) jl[l] = J[l < k ? l : 0];
  for (int i = lane; i < n && map_rows; i += 64) {
    if (cid[i] == cl) continue;
    double* Ei = Eb + (long long)i * n;
    double old[RR_SMALL];
#pragma unroll
    for (int l = 0; l < RR_SMALL; l++) old[l] = l < k ? Ei[jl[l]] : 0.0;
#pragma unroll
    for (int c = 0; c < RR_SMALL; c++) {
      if (c >= k) break;
      double v = 0.0;
#pragma unroll
      for (int l = 0; l < RR_SMALL; l++) v += old[l] * T[l * 8 + c];
      Ei[jl[c]] = v;
    }
  }
  if (in) Eb[(long long)ja * n + jb] = g - (a == b ? 1.0 : 0.0);
}

// The OA step as four launches over the active blocks:
//   k_oa_sym      32x32 tile pairs (i,j)/(j,i): S and Rp, computed on and above the diagonal only,
//                 completed by mirroring (exactly symmetric S and Rp keep E + E^T = -(Rp - I) exact,
//                 so the step keeps orthogonality even where a small gap magnifies GEMM noise); per
//                 block the largest off-diagonal
//                 rotation residual |rho_ij| = |S_ij - (lt_i + lt_j)/2 Rp_ij| (the orthogonality
//                 defect, which needs no gap, is left out) and the largest |Rp_ij - delta_ij|
//   k_oa_prep     one workgroup per block: lt, delta, ascending sort, cluster ids, and the rounding
//                 noise level sqrt(n) eps max|lt| of the entries of S
//   k_oa_entries  32x32 tiles: E_ij (coalesced), largest out-of-cluster residual |rho_ij|
// conv[4b..4b+3] = (largest residual, largest orthogonality defect, noise level, largest out-of-cluster
// rotation |rho_ij / (lt_j - lt_i)|) of active block b.
//   k_oa_rr       one workgroup per block: Rayleigh-Ritz of every cluster of two or more
constexpr int OT = 32;
constexpr int OA_CONV = 4;  // entries of conv per block

// A block takes its last step when its out-of-cluster residuals are within e_tol times the rounding
// noise of S and its orthogonality defect is at most d_tol.  Its out-of-cluster corrections are then
// rounding noise over a gap, which could only spoil orthogonality: the last step applies the cluster
// rotations and the normalisation alone.
__device__ __forceinline__ bool oa_converged(const double* conv, int b, double e_tol, double d_tol) {
  return conv[OA_CONV * b] <= e_tol * conv[OA_CONV * b + 2] && conv[OA_CONV * b + 1] <= d_tol;
}
// A block leaves the batch after its last step, or after a full step taken from residuals within
// e_pred times the noise, an orthogonality defect of at most d_pred and out-of-cluster rotations r
// with n r^2 <= d_tol: that step's update lands at the noise (the residuals fall quadratically, to
// rho^2 / gap with every out-of-cluster gap at least 1e5 noise, so e_pred = 300 leaves below one
// noise unit; the normalisation defect squares, 1e-6 -> 1e-12; X + XE is orthogonal only to first
// order in the rotations, so they must be small enough that their second-order loss stays within
// d_tol -- a pair just outside a cluster can take a rotation of 1e-3), so the verifying step that
// would follow is skipped.
__device__ __host__ __forceinline__ bool oa_predicted(const double* conv, int b, int n, double d_tol, double e_pred,
                                                      double d_pred) {
  const double r = conv[OA_CONV * b + 3];
  return conv[OA_CONV * b] <= e_pred * conv[OA_CONV * b + 2] && conv[OA_CONV * b + 1] <= d_pred &&
         (double)n * r * r <= d_tol;
}
__device__ __forceinline__ bool oa_done(const double* conv, int b, int n, double e_tol, double d_tol, double e_pred,
                                        double d_pred) {
  return oa_converged(conv, b, e_tol, d_tol) || oa_predicted(conv, b, n, d_tol, e_pred, d_pred);
}

__global__ void __launch_bounds__(256) k_oa_mask(const EigBlock* __restrict__ blocks, const int4* __restrict__ tiles,
                                                 const int* __restrict__ cid_g, const double* __restrict__ conv,
                                                 double e_tol, double d_tol, double* __restrict__ E) {
  const int4 t = tiles[blockIdx.x];
  if (!oa_converged(conv, t.x, e_tol, d_tol)) return;
  const EigBlock blk = blocks[t.x];
  const int n = blk.n;
  double* Eb = E + blk.off;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int j = t.z * OT + tx;
  if (j >= n) return;
  const int cj = cid_g[blk.e_off + j];
  for (int rr = ty; rr < OT; rr += 8) {
    const int i = t.y * OT + rr;
    if (i >= n) break;
    if (cid_g[blk.e_off + i] != cj) Eb[(long long)i * n + j] = 0.0;
  }
}


__global__ void __launch_bounds__(256) k_oa_sym(const EigBlock* __restrict__ blocks, const int4* __restrict__ tiles,
                                                double* __restrict__ S, double* __restrict__ Rp,
                                                double* __restrict__ offmax, double* __restrict__ conv) {
  __shared__ double tS[OT][OT + 1], tR[OT][OT + 1];
  __shared__ double red[8];
  const int4 t = tiles[blockIdx.x];
  const EigBlock blk = blocks[t.x];
  const int n = blk.n;
  double* Sb = S + blk.off;
  double* Rb = Rp + blk.off;
  const int r0 = t.y * OT, c0 = t.z * OT;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  // the mirror tile (rows c0.., columns r0..)
  for (int rr = ty; rr < OT; rr += 8) {
    const int gr = c0 + rr, gc = r0 + tx;
    if (gr < n && gc < n) {
      tS[rr][tx] = Sb[(long long)gr * n + gc];
      tR[rr][tx] = Rb[(long long)gr * n + gc];
    }
  }
  double sv[4], rv[4];
  const int j = c0 + tx;
  double ltj = 0.0;
  if (j < n) ltj = Sb[(long long)j * n + j] / Rb[(long long)j * n + j];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = r0 + ty + 8 * q;
    if (i < n && j < n) {
      sv[q] = Sb[(long long)i * n + j];
      rv[q] = Rb[(long long)i * n + j];
    }
  }
  __syncthreads();  // every own element read before any mirror element is written
  double mx = 0.0, dm = 0.0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int rr = ty + 8 * q, i = r0 + rr;
    if (i >= n || j >= n) continue;
    // S and Rp are computed on and above the diagonal only: take the upper element for both
    const double s = i <= j ? sv[q] : tS[tx][rr];
    const double r = i <= j ? rv[q] : tR[tx][rr];
    if (i != j) {
      const double lti = Sb[(long long)i * n + i] / Rb[(long long)i * n + i];
      mx = fmax(mx, fabs(s - 0.5 * (lti + ltj) * r));
      dm = fmax(dm, fabs(r));
    } else {
      dm = fmax(dm, fabs(1.0 - r));
    }
    Sb[(long long)i * n + j] = s;
    Sb[(long long)j * n + i] = s;
    Rb[(long long)i * n + j] = r;
    Rb[(long long)j * n + i] = r;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, o, 64));
    dm = fmax(dm, __shfl_xor(dm, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = mx;
    red[4 + (threadIdx.x >> 6)] = dm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    dm = fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
    // non-negative doubles order like their bit patterns
    atomicMax((unsigned long long*)&offmax[t.x], (unsigned long long)__double_as_longlong(mx));
    atomicMax((unsigned long long*)&conv[OA_CONV * t.x + 1], (unsigned long long)__double_as_longlong(dm));
  }
}

// lt_i = S_ii / Rp_ii; delta = max(32 max|rho| + 64 eps max|lt|, 1e5 sqrt(n) eps max|lt|), so that no
// out-of-cluster rotation exceeds 1/32; below the
// rounding noise of S_ij (about sqrt(n) eps max|lt|) the OA quotient is noise over gap, so pairs
// closer than 1e5 times that noise are handled as clusters, which caps the out-of-cluster corrections
// made of pure rounding at about 1e-5 and keeps the iteration at its fixed point once converged.
// Clusters are the runs of sorted lt with gaps <= delta; cid = sorted position of the cluster's first
// member, sidx = sorted order.
__global__ void __launch_bounds__(1024) k_oa_prep(const EigBlock* __restrict__ blocks, const double* __restrict__ S,
                                                  const double* __restrict__ Rp, const double* __restrict__ offmax,
                                                  double* __restrict__ lam, int* __restrict__ cid_g,
                                                  int* __restrict__ sidx_g, int* __restrict__ stats,
                                                  double* __restrict__ conv) {
  extern __shared__ double sh[];
  const EigBlock blk = blocks[blockIdx.x];
  const int n = blk.n;
  if (n == 0) return;
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  double* key = sh;              // [np2]
  int* idx = (int*)(key + np2);  // [np2]
  int* head = idx + np2;         // [n] 1 where a cluster starts (sorted position)
  __shared__ double red[16];
  __shared__ int s_maxk;
  const int tid = threadIdx.x;
  if (tid == 0) s_maxk = 1;
  const double* Sb = S + blk.off;
  const double* Rb = Rp + blk.off;
  double ml = 0.0;
  for (int i = tid; i < np2; i += blockDim.x) {
    double v = 1e300;
    if (i < n) {
      v = Sb[(long long)i * n + i] / Rb[(long long)i * n + i];
      lam[blk.e_off + i] = v;
      ml = fmax(ml, fabs(v));
    }
    key[i] = v;
    idx[i] = i;
  }
  const double bq = block_max(ml, red);
  for (int k = 2; k <= np2; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = tid; i < np2; i += blockDim.x) {
        const int l = i ^ jj;
        if (l > i) {
          const bool up = ((i & k) == 0);
          const double a = key[i], bb = key[l];
          if ((a > bb) == up) {
            key[i] = bb;
            key[l] = a;
            const int t = idx[i];
            idx[i] = idx[l];
            idx[l] = t;
          }
        }
      }
      __syncthreads();
    }
  const double eps = 2.220446049250313e-16;
  const double noise = sqrt((double)n) * eps * bq;
  const double delta = fmax(32.0 * offmax[blockIdx.x] + 64.0 * eps * bq, 1e5 * noise);
  if (tid == 0) conv[OA_CONV * blockIdx.x + 2] = noise;
  for (int s = tid; s < n; s += blockDim.x) head[s] = (s == 0) || (key[s] - key[s - 1] > delta);
  __syncthreads();
  // cluster id = sorted position of the cluster's first member (ids are only compared), and the
  // widest cluster, all positions in parallel
  for (int s = tid; s < n; s += blockDim.x) {
    int s0 = s;
    while (!head[s0]) s0--;  // head[0] is set
    cid_g[blk.e_off + idx[s]] = s0;
    sidx_g[blk.e_off + s] = idx[s];
    if (head[s]) {
      int e = s + 1;
      while (e < n && !head[e]) e++;
      atomicMax(&s_maxk, e - s);
    }
  }
  __syncthreads();
  if (tid == 0) {
    atomicMax(&stats[0], s_maxk);
    if (s_maxk > RR_MAX) atomicOr(&stats[1], 1);
  }
}

__global__ void __launch_bounds__(256) k_oa_entries(const EigBlock* __restrict__ blocks, const int4* __restrict__ tiles,
                                                    const double* __restrict__ S, const double* __restrict__ Rp,
                                                    const double* __restrict__ lam, const int* __restrict__ cid_g,
                                                    double* __restrict__ E, double* __restrict__ conv) {
  __shared__ double red[8];
  const int4 t = tiles[blockIdx.x];
  const EigBlock blk = blocks[t.x];
  const int n = blk.n;
  const double* Sb = S + blk.off;
  const double* Rb = Rp + blk.off;
  double* Eb = E + blk.off;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int j = t.z * OT + tx;
  double em = 0.0, rm = 0.0;  // largest residual and largest rotation of the tile
  if (j < n) {
    const double ltj = lam[blk.e_off + j];
    const int cj = cid_g[blk.e_off + j];
    for (int rr = ty; rr < OT; rr += 8) {
      const int i = t.y * OT + rr;
      if (i >= n) break;
      const long long e = (long long)i * n + j;
      double v;
      if (i == j) {
        v = 0.5 * (1.0 - Rb[e]);
      } else if (cid_g[blk.e_off + i] == cj) {
        v = 0.0;
      } else {
        // E_ij = (S_ij - lt_j R_ij)/(lt_j - lt_i), split into its rotation rho_ij/(lt_j - lt_i),
        // rho_ij = S_ij - (lt_i + lt_j)/2 R_ij, and its normalisation part -R_ij/2: the split keeps
        // E + E^T = -(R - I) exact in floating point and avoids a quotient of two small numbers
        const double lti = lam[blk.e_off + i];
        const double rho = Sb[e] - 0.5 * (lti + ltj) * Rb[e];
        const double rot = rho / (ltj - lti);
        v = rot - 0.5 * Rb[e];
        em = fmax(em, fabs(rho));
        rm = fmax(rm, fabs(rot));
      }
      Eb[e] = v;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    em = fmax(em, __shfl_xor(em, o, 64));
    rm = fmax(rm, __shfl_xor(rm, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = em;
    red[4 + (threadIdx.x >> 6)] = rm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    em = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    rm = fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
    // non-negative doubles order as their bit patterns
    atomicMax((unsigned long long*)&conv[OA_CONV * t.x], (unsigned long long)__double_as_longlong(em));
    atomicMax((unsigned long long*)&conv[OA_CONV * t.x + 3], (unsigned long long)__double_as_longlong(rm));
  }
}

// Rayleigh-Ritz on every cluster of two or more: M = 1.5 I - 0.5 Rp_JJ, (M S_JJ M) Q = Q mu,
// E_JJ = M Q - I and the out-of-cluster rows of E(:,J) mapped through M Q.  Clusters of up to
// RR_SMALL members one wave each, the wider ones in turn by the whole workgroup.
__global__ void __launch_bounds__(1024) k_oa_rr(const EigBlock* __restrict__ blocks, const double* __restrict__ S,
                                                const double* __restrict__ Rp, double* __restrict__ E,
                                                double* __restrict__ lam, const int* __restrict__ cid_g,
                                                const int* __restrict__ sidx_g, double* __restrict__ scrW,
                                                double* __restrict__ scrG, double* __restrict__ scrT,
                                                double* __restrict__ rrv, const double* __restrict__ conv,
                                                double e_tol, double d_tol) {
  extern __shared__ double sh[];
  const EigBlock blk = blocks[blockIdx.x];
  const int n = blk.n;
  if (n < 2) return;
  const bool last = oa_converged(conv, blockIdx.x, e_tol, d_tol);
  const double* Sb = S + blk.off;
  const double* Rb = Rp + blk.off;
  double* Eb = E + blk.off;
  const int* cidb = cid_g + blk.e_off;
  const int* idx = sidx_g + blk.e_off;
  double* rrW = sh;                      // [RR_MAX*RR_MAX]
  double* rrM = rrW + RR_MAX * RR_MAX;   // [RR_MAX*RR_MAX] M, then G = M Q
  double* rrMu = rrM + RR_MAX * RR_MAX;  // [RR_MAX]
  double* rrN = rrMu + RR_MAX;           // [RR_MAX]
  double* red = rrN + RR_MAX;            // [32]
  int* rrP = (int*)(red + 32);           // [RR_MAX]
  int* cstart = rrP + RR_MAX;            // [n + 2] cluster (start, end) of width >= 2 (sorted positions)
  __shared__ int s_nc;
  const int tid = threadIdx.x;
  // clusters are runs of equal id in sorted order: ids staged in LDS, every run start found in
  // parallel (the list order does not matter, clusters are disjoint)
  int* csort = cstart + n + 2;  // [n] cluster id by sorted position
  if (tid == 0) s_nc = 0;
  for (int s = tid; s < n; s += blockDim.x) csort[s] = cidb[idx[s]];
  __syncthreads();
  for (int s = tid; s < n; s += blockDim.x) {
    if (s > 0 && csort[s] == csort[s - 1]) continue;
    int e = s + 1;
    while (e < n && csort[e] == csort[s]) e++;
    if (e - s >= 2) {
      const int slot = atomicAdd(&s_nc, 1);
      cstart[2 * slot] = s;
      cstart[2 * slot + 1] = e;
    }
  }
  __syncthreads();
  const int ncl = s_nc;
  // clusters of 2..RR_SMALL members: one wave each, concurrently (they touch disjoint columns of E)
  {
    double* wscr = (double*)(((uintptr_t)(csort + n + 2) + 7) & ~(uintptr_t)7);
    const int wave = tid >> 6;
    for (int c = wave; c < ncl; c += blockDim.x / 64) {
      const int s0 = cstart[2 * c], k = cstart[2 * c + 1] - s0;
      if (k <= RR_SMALL)
        rr_small_wave(Sb, Rb, Eb, n, k, idx + s0, cidb, cidb[idx[s0]], lam + blk.e_off, wscr + 192 * wave, !last);
    }
    __syncthreads();
  }
  auto Ss = [&](int i, int j) { return Sb[(long long)i * n + j]; };
  auto Rs = [&](int i, int j) { return Rb[(long long)i * n + j]; };
  for (int c = 0; c < ncl; c++) {
    const int s0 = cstart[2 * c], k = cstart[2 * c + 1] - s0;
    if (k <= RR_SMALL) continue;
    const int cl = cidb[idx[s0]];
    if (k > RR_MAX) {
      rr_wide_cluster(Sb, Rb, Eb, n, k, idx + s0, cidb, cl, scrW + blk.off, scrG + blk.off, scrT + blk.off,
                      rrv + 3 * (long long)blk.e_off, lam + blk.e_off, red, !last);
      continue;
    }
    // M = 1.5 I - 0.5 Rp_JJ;  T = M S_JJ M  (row-major k x k in rrW after two products)
    for (int e = tid; e < k * k; e += blockDim.x) {
      const int a = e / k, bq = e % k;
      rrM[e] = (a == bq ? 1.5 : 0.0) - 0.5 * Rs(idx[s0 + a], idx[s0 + bq]);
    }
    __syncthreads();
    for (int e = tid; e < k * k; e += blockDim.x) {  // rrW = S_JJ M
      const int a = e / k, bq = e % k;
      double s = 0;
      for (int l = 0; l < k; l++) s += Ss(idx[s0 + a], idx[s0 + l]) * rrM[l * k + bq];
      rrW[e] = s;
    }
    __syncthreads();
    // T = M (S M), held in registers until every thread has read rrW
    double tv[RR_MAX * RR_MAX / 1024 + 1];
    int cnt = 0;
    for (int e = tid; e < k * k; e += blockDim.x) {
      const int a = e / k, bq = e % k;
      double s = 0;
      for (int l = 0; l < k; l++) s += rrM[a * k + l] * rrW[l * k + bq];
      tv[cnt++] = s;
    }
    __syncthreads();
    cnt = 0;
    for (int e = tid; e < k * k; e += blockDim.x) rrW[e] = tv[cnt++];
    __syncthreads();
    // shift by the cluster centre and make it positive definite: T' = T - c I + sig I
    if (tid == 0) {
      double cen = 0;
      for (int a = 0; a < k; a++) cen += rrW[a * k + a];
      cen /= k;
      double rad = 0;
      for (int a = 0; a < k; a++) {
        double s = 0;
        for (int bq = 0; bq < k; bq++) s += fabs(rrW[a * k + bq] - (a == bq ? cen : 0.0));
        rad = fmax(rad, s);
      }
      // an exactly degenerate cluster has rad == 0: any positive shift then gives Q = I exactly;
      // the floor keeps the squared column norms far from underflow
      const double sig = rad > 1e-30 * fmax(1.0, fabs(cen)) ? 1.05 * rad : 1.0;
      // (T_aa - cen) is exact for a tight cluster; adding sig afterwards keeps the shift even when it
      // is far below one ulp of cen (T_aa + (sig - cen) would round it away)
      for (int a = 0; a < k; a++) rrW[a * k + a] = (rrW[a * k + a] - cen) + sig;
      red[0] = cen;
      red[1] = sig;
    }
    __syncthreads();
    rr_jacobi(rrW, k, rrMu, rrP, rrN);  // symmetric: column-major == row-major
    const double cen = red[0], sig = red[1];
    // Q(:,c) = W(:,perm[c]) / nrm ;  G = M Q
    cnt = 0;
    for (int e = tid; e < k * k; e += blockDim.x) {
      const int a = e / k, cc = e % k;
      const int src = rrP[cc];
      double s = 0;
      for (int l = 0; l < k; l++) s += rrM[a * k + l] * (rrW[src * k + l] / rrN[src]);
      tv[cnt++] = s;
    }
    __syncthreads();
    cnt = 0;
    for (int e = tid; e < k * k; e += blockDim.x) rrM[e] = tv[cnt++];  // rrM := G
    __syncthreads();
    // out-of-cluster rows: E(i, J) <- E(i, J) G.  One thread owns a whole row i, so every old value
    // of the row is read before any of them is overwritten.  (On a block's last step k_oa_mask
    // clears these entries instead.)
    for (int i = tid; i < n && !last; i += blockDim.x) {
      if (cidb[i] == cl) continue;
      double old[RR_MAX];
      for (int l = 0; l < k; l++) old[l] = Eb[(long long)i * n + idx[s0 + l]];
      for (int cc = 0; cc < k; cc++) {
        double s = 0;
        for (int l = 0; l < k; l++) s += old[l] * rrM[l * k + cc];
        Eb[(long long)i * n + idx[s0 + cc]] = s;
      }
    }
    __syncthreads();
    for (int e = tid; e < k * k; e += blockDim.x) {
      const int a = e / k, cc = e % k;
      Eb[(long long)idx[s0 + a] * n + idx[s0 + cc]] = rrM[e] - (a == cc ? 1.0 : 0.0);
    }
    for (int cc = tid; cc < k; cc += blockDim.x) lam[blk.e_off + idx[s0 + cc]] = rrMu[cc] - sig + cen;
    __syncthreads();
  }
}

// Blocks that took their last refinement step leave the batch with their vectors in X; this copies
// them to the other iterate buffer as well (one launch for all of them: grid = blocks x 8).
__global__ void __launch_bounds__(256) k_copy_converged(const EigBlock* __restrict__ blocks,
                                                        const double* __restrict__ conv, double e_tol, double d_tol,
                                                        double e_pred, double d_pred, const double* __restrict__ src,
                                                        double* __restrict__ dst) {
  const int a = blockIdx.x >> 3, part = blockIdx.x & 7;
  const EigBlock blk = blocks[a];
  if (!oa_done(conv, a, blk.n, e_tol, d_tol, e_pred, d_pred)) return;
  const long long nn = (long long)blk.n * blk.n;
  for (long long e = part * 256LL + threadIdx.x; e < nn; e += 8 * 256LL) dst[blk.off + e] = src[blk.off + e];
}

void launch_copy_converged(const EigBlock* blocks, int nb, const double* conv, double e_tol, double d_tol,
                           double e_pred, double d_pred, const double* src, double* dst, hipStream_t s) {
  if (nb <= 0) return;
  NRGK_LAUNCH(k_copy_converged, dim3(nb * 8), dim3(256), 0, s, blocks, conv, e_tol, d_tol, e_pred, d_pred, src,
                     dst);
  NRGK_HIP(hipGetLastError());
}

// Final ordering: U(:, c) = X(:, perm[c]) with eigenvalues ascending (dsyevd order).
__global__ void __launch_bounds__(1024) k_final_sort(const EigBlock* __restrict__ blocks,
                                                     const double* __restrict__ lam_in, double* __restrict__ lam_out,
                                                     int* __restrict__ perm) {
  extern __shared__ double sh[];
  const EigBlock blk = blocks[blockIdx.x];
  const int n = blk.n;
  if (n == 0) return;
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  double* key = sh;
  int* idx = (int*)(sh + np2);
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    key[i] = i < n ? lam_in[blk.e_off + i] : 1e300;
    idx[i] = i;
  }
  __syncthreads();
  for (int k = 2; k <= np2; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = threadIdx.x; i < np2; i += blockDim.x) {
        const int l = i ^ jj;
        if (l > i) {
          const bool up = ((i & k) == 0);
          const double a = key[i], bb = key[l];
          if ((a > bb) == up) {
            key[i] = bb;
            key[l] = a;
            const int t = idx[i];
            idx[i] = idx[l];
            idx[l] = t;
          }
        }
      }
      __syncthreads();
    }
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    lam_out[blk.e_off + c] = key[c];
    perm[blk.e_off + c] = idx[c];
  }
}

// U(r, c) = X(r, perm c), one 32 x 32 tile of one block per workgroup
__global__ void __launch_bounds__(256) k_gather_u(const EigBlock* __restrict__ blocks, const int4* __restrict__ tiles,
                                                  const double* __restrict__ X, const int* __restrict__ perm,
                                                  double* __restrict__ U) {
  const int4 t = tiles[blockIdx.x];
  const EigBlock blk = blocks[t.x];
  const int n = blk.n, r0 = 32 * t.y, c0 = 32 * t.z;
  const double* Xb = X + blk.off;
  double* Ub = U + blk.off;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = threadIdx.x + 256 * k, rr = e >> 5, c = e & 31;
    if (c0 + c < n && r0 + rr < n) {
      const long long row = (long long)(r0 + rr) * n;
      Ub[row + c0 + c] = Xb[row + perm[blk.e_off + c0 + c]];
    }
  }
}

size_t normalize_sort_lds(int nmax) {
  int np2 = 1;
  while (np2 < nmax) np2 <<= 1;
  return (size_t)(np2 + (np2 + 1) / 2 + 1 + nmax) * sizeof(double) + 64;
}
size_t oa_prep_lds(int nmax) {
  int np2 = 1;
  while (np2 < nmax) np2 <<= 1;
  return (size_t)np2 * 12 + (size_t)nmax * 4 + 64;
}
size_t oa_rr_lds(int nmax) {
  return (size_t)(2 * RR_MAX * RR_MAX + 2 * RR_MAX + 32) * 8 + (size_t)RR_MAX * 4 + (size_t)(nmax + 2) * 8 + 64 +
         16 * 192 * sizeof(double);
}
size_t final_sort_lds(int nmax) {
  int np2 = 1;
  while (np2 < nmax) np2 <<= 1;
  return (size_t)np2 * 12 + 64;
}

void launch_jacobi_round(const JacTask* tasks, int n_tasks, float* W, int round, int* flags, double* dgram, int* dgv,
                         float tol, float ang_tol, int max_inner, double jac_far, hipStream_t s) {
  if (n_tasks <= 0) return;
  NRGK_LAUNCH(k_jacobi_round, dim3(n_tasks), dim3(JAC_THREADS), JAC_W * JAC_CS * sizeof(float), s, tasks, W,
                     round, flags, dgram, dgv, tol, ang_tol, max_inner, jac_far * jac_far);
  NRGK_HIP(hipGetLastError());
}
void launch_normalize_sort(const EigBlock* blocks, int nb, const int4* tiles, int n_tiles, const int4* groups,
                           int n_groups, const float* W, const float* sigma, double* X, double* lam, int* perm,
                           double* nrm_s, double* nrm_g, size_t lds, hipStream_t s) {
  if (nb <= 0) return;
  NRGK_LAUNCH(k_col_norms, dim3(n_groups), dim3(256), 0, s, blocks, groups, W, nrm_g);
  NRGK_HIP(hipGetLastError());
  NRGK_LAUNCH(k_normalize_sort, dim3(nb), dim3(1024), lds, s, blocks, nrm_g, sigma, lam, perm, nrm_s);
  NRGK_HIP(hipGetLastError());
  if (n_tiles <= 0) return;
  NRGK_LAUNCH(k_gather_x, dim3(n_tiles), dim3(256), 0, s, blocks, tiles, W, perm, nrm_s, X);
  NRGK_HIP(hipGetLastError());
}
void launch_oa(const EigBlock* blocks, int nb, const int4* sym_tiles, int n_sym, const int4* all_tiles, int n_all,
               double* S, double* Rp, double* E, double* lam, int* cid, int* sidx, double* offmax, int* stats,
               double* conv, double* scrW, double* scrG, double* scrT, double* rrv, int nmax, double e_tol,
               double d_tol, hipStream_t s) {
  if (nb <= 0) return;
  NRGK_LAUNCH(k_oa_sym, dim3(n_sym), dim3(256), 0, s, blocks, sym_tiles, S, Rp, offmax, conv);
  NRGK_HIP(hipGetLastError());
  NRGK_LAUNCH(k_oa_prep, dim3(nb), dim3(1024), oa_prep_lds(nmax), s, blocks, S, Rp, offmax, lam, cid, sidx,
                     stats, conv);
  NRGK_HIP(hipGetLastError());
  NRGK_LAUNCH(k_oa_entries, dim3(n_all), dim3(256), 0, s, blocks, all_tiles, S, Rp, lam, cid, E, conv);
  NRGK_HIP(hipGetLastError());
  NRGK_LAUNCH(k_oa_rr, dim3(nb), dim3(1024), oa_rr_lds(nmax), s, blocks, S, Rp, E, lam, cid, sidx, scrW, scrG,
                     scrT, rrv, conv, e_tol, d_tol);
  NRGK_HIP(hipGetLastError());
  NRGK_LAUNCH(k_oa_mask, dim3(n_all), dim3(256), 0, s, blocks, all_tiles, cid, conv, e_tol, d_tol, E);
  NRGK_HIP(hipGetLastError());
}
void launch_final_sort(const EigBlock* blocks, int nb, const int4* tiles, int n_tiles, const double* X,
                       const double* lam_in, double* U, double* lam_out, int* perm, size_t lds, hipStream_t s) {
  if (nb <= 0) return;
  NRGK_LAUNCH(k_final_sort, dim3(nb), dim3(1024), lds, s, blocks, lam_in, lam_out, perm);
  NRGK_HIP(hipGetLastError());
  if (n_tiles <= 0) return;
  NRGK_LAUNCH(k_gather_u, dim3(n_tiles), dim3(256), 0, s, blocks, tiles, X, perm, U);
  NRGK_HIP(hipGetLastError());
}
void set_kernel_lds_limits(int bytes) {
  set_trunc_lds_limit();
  NRGK_HIP(hipFuncSetAttribute((const void*)k_jacobi_round, hipFuncAttributeMaxDynamicSharedMemorySize,
                                JAC_W * JAC_CS * sizeof(float)));
  NRGK_HIP(hipFuncSetAttribute((const void*)k_normalize_sort, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  NRGK_HIP(hipFuncSetAttribute((const void*)k_oa_prep, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  NRGK_HIP(hipFuncSetAttribute((const void*)k_oa_rr, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  NRGK_HIP(hipFuncSetAttribute((const void*)k_final_sort, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
}

}  // namespace nrgk
