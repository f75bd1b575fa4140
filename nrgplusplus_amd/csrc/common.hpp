// Shared host/device definitions of libnrgk.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nrgk.h"

namespace nrgk {

// ------------------------------------------------------------------ errors
// Host code throws nrgk::Error; the C entry points catch it, keep the message for
// nrgk_last_error() and return the status.
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define NRGK_HIP(call)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::nrgk::Error(NRGK_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                            " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define NRGK_REQUIRE(cond, status, msg)            \
  do {                                             \
    if (!(cond)) throw ::nrgk::Error(status, msg); \
  } while (0)

void set_last_error(const std::string& m);

// Kernel launches (the library's own kernels; a rocBLAS or rocSOLVER call counts once) and host
// synchronisations made by the calling thread.  A chain is driven by one thread at a time, so the
// difference of two readings around a step is that step's count.
struct ThreadCounts {
  long long launches = 0, syncs = 0;
};
ThreadCounts& thread_counts();
#define NRGK_LAUNCH(...)                    \
  do {                                      \
    ::nrgk::thread_counts().launches++;     \
    hipLaunchKernelGGL(__VA_ARGS__);        \
  } while (0)
// hipStreamSynchronize, counted
inline hipError_t stream_sync(hipStream_t s) {
  thread_counts().syncs++;
  return hipStreamSynchronize(s);
}

// ------------------------------------------------------------------ device buffers
// A device allocation that only grows; per-iteration scratch reuses it without hipMalloc in the
// steady state.  Every allocation is charged to the owning context's capacity (HBM admission).
struct Ctx;
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  Ctx* ctx = nullptr;
  DevBuf() = default;
  explicit DevBuf(Ctx* c) : ctx(c) {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), cap(o.cap), ctx(o.ctx) {
    o.p = nullptr;
    o.cap = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      cap = o.cap;
      ctx = o.ctx;
      o.p = nullptr;
      o.cap = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void ensure(size_t bytes);  // grows (contents not preserved)
  void release();
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Ctx {
  int device = 0;
  size_t capacity = 0;
  std::atomic<size_t> used{0};  // chains of one context may be stepped from several host threads
  int lds_max = 65536;  // bytes of LDS one workgroup may allocate
  int n_cu = 0;
  std::string arch;     // gcnArchName of the device
  // the eigensolver crossover: blocks wider than this leave the batched Jacobi for the Householder
  // path (768, the widest LDS-resident Jacobi panel, unless the autotuner's cache says otherwise)
  int mixed_max = 768;
  bool tuned = false;
  void charge(size_t bytes) {
    size_t u = used.load();
    do {
      if (capacity && u + bytes > capacity)
        throw Error(NRGK_ERR_NOMEM, "nrgk: device capacity exceeded (" + std::to_string(u + bytes) + " > " +
                                        std::to_string(capacity) + " bytes)");
    } while (!used.compare_exchange_weak(u, u + bytes));
  }
  void refund(size_t bytes) {
    size_t u = used.load();
    while (!used.compare_exchange_weak(u, bytes < u ? u - bytes : 0)) {
    }
  }
};

// Page-locked staging for one stream's small transfers.  put() copies host data into pinned memory
// and enqueues one asynchronous copy to its device twin, returning the device address; get()
// enqueues a device-to-host copy into pinned memory and returns the host address, readable once the
// stream is synchronised.  Regions are handed out front to back and stay valid until reset(), which
// the owner calls after synchronising the stream.  Running out of room opens a further chunk (no
// region moves while in use); reset() merges the chunks into one.
struct Staging {
  struct Chunk {
    char* host = nullptr;
    char* dev = nullptr;
    size_t cap = 0, used = 0;
  };
  std::vector<Chunk> chunks;
  Ctx* ctx = nullptr;
  Staging() = default;
  Staging(const Staging&) = delete;
  Staging& operator=(const Staging&) = delete;
  ~Staging() { release(); }
  void release() {
    for (auto& c : chunks) {
      if (c.host) (void)hipHostFree(c.host);
      if (c.dev) {
        (void)hipFree(c.dev);
        if (ctx) ctx->refund(c.cap);
      }
    }
    chunks.clear();
  }
  void add_chunk(size_t cap) {
    Chunk c;
    c.cap = cap;
    if (ctx) ctx->charge(cap);
    hipError_t e = hipHostMalloc((void**)&c.host, cap, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void**)&c.dev, cap);
    if (e != hipSuccess) {
      if (c.host) (void)hipHostFree(c.host);
      if (ctx) ctx->refund(cap);
      throw Error(NRGK_ERR_NOMEM, std::string("nrgk: staging allocation (") + std::to_string(cap) + "): " +
                                      hipGetErrorString(e));
    }
    chunks.push_back(c);
  }
  // an aligned region of `bytes`: (host, device) addresses
  std::pair<char*, char*> take(size_t bytes) {
    const size_t need = (bytes + 255) & ~size_t(255);
    if (chunks.empty() || chunks.back().used + need > chunks.back().cap)
      add_chunk(std::max<size_t>({need, chunks.empty() ? 0 : 2 * chunks.back().cap, size_t(1) << 20}));
    Chunk& c = chunks.back();
    const size_t o = c.used;
    c.used += need;
    return {c.host + o, c.dev + o};
  }
  template <class T>
  T* put(const T* src, size_t n, hipStream_t s) {
    auto [h, d] = take(std::max<size_t>(1, n) * sizeof(T));
    if (n) {
      std::memcpy(h, src, n * sizeof(T));
      NRGK_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    return reinterpret_cast<T*>(d);
  }
  template <class T>
  T* put(const std::vector<T>& v, hipStream_t s) {
    return put(v.data(), v.size(), s);
  }
  template <class T>
  T* put_zero(size_t n, hipStream_t s) {
    auto [h, d] = take(std::max<size_t>(1, n) * sizeof(T));
    std::memset(h, 0, std::max<size_t>(1, n) * sizeof(T));
    NRGK_HIP(hipMemcpyAsync(d, h, std::max<size_t>(1, n) * sizeof(T), hipMemcpyHostToDevice, s));
    return reinterpret_cast<T*>(d);
  }
  // Several arrays sent with one copy: open() a region for their total span() bytes, place() each
  // array (host side only; returns its device address), then commit() enqueues a single
  // host-to-device copy of the region.
  struct Region {
    char* h = nullptr;
    char* d = nullptr;
    size_t used = 0, cap = 0;
  };
  template <class T>
  static size_t span(size_t n) {
    return (std::max<size_t>(1, n) * sizeof(T) + 255) & ~size_t(255);
  }
  Region open(size_t bytes) {
    auto [h, d] = take(bytes);
    return Region{h, d, 0, (bytes + 255) & ~size_t(255)};
  }
  template <class T>
  T* place(Region& r, const T* src, size_t n) {
    const size_t sz = span<T>(n);
    if (r.used + sz > r.cap) throw Error(NRGK_ERR_STATE, "nrgk: staging region overflow");
    if (n) std::memcpy(r.h + r.used, src, n * sizeof(T));
    T* d = reinterpret_cast<T*>(r.d + r.used);
    r.used += sz;
    return d;
  }
  template <class T>
  T* place(Region& r, const std::vector<T>& v) {
    return place(r, v.data(), v.size());
  }
  template <class T>
  T* place_zero(Region& r, size_t n) {
    const size_t sz = span<T>(n);
    if (r.used + sz > r.cap) throw Error(NRGK_ERR_STATE, "nrgk: staging region overflow");
    std::memset(r.h + r.used, 0, sz);
    T* d = reinterpret_cast<T*>(r.d + r.used);
    r.used += sz;
    return d;
  }
  void commit(const Region& r, hipStream_t s) {
    if (r.used) NRGK_HIP(hipMemcpyAsync(r.d, r.h, r.used, hipMemcpyHostToDevice, s));
  }
  template <class T>
  const T* get(const T* dsrc, size_t n, hipStream_t s) {
    auto [h, d] = take(std::max<size_t>(1, n) * sizeof(T));
    (void)d;
    if (n) NRGK_HIP(hipMemcpyAsync(h, dsrc, n * sizeof(T), hipMemcpyDeviceToHost, s));
    return reinterpret_cast<const T*>(h);
  }
  void reset() {
    if (chunks.size() > 1) {
      size_t total = 0;
      for (auto& c : chunks) total += c.cap;
      release();
      add_chunk(total);
    }
    if (!chunks.empty()) chunks.back().used = 0;
  }
};

template <class T>
inline void h2d(T* dst, const T* src, size_t n, hipStream_t s) {
  if (n) NRGK_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T>
inline void d2h(T* dst, const T* src, size_t n, hipStream_t s) {
  if (n) NRGK_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

// ------------------------------------------------------------------ descriptors (host <-> device)
// Sector assembly (nrgcore.hpp:146-230).  One AsmSub per coupled (previous sector ii, bath sector
// bb) of a new sector, in coupled_nQ_index order; one AsmCon per (sub-block k, sub-block kp, op p)
// for which both F_prev[p](ii_k, ii_kp) and F_bath[p](bb_kp, bb_k) exist, in the reference's
// program order (k outer, kp inner, p innermost).
struct AsmSub {
  int kidx;   // first row of the sub-block inside the sector block
  int kii;    // kept count of the previous sector ii
  int dbb;    // dimension of bath sector bb
  int e_off;  // offset of E_prev[ii] in the kept-energy array
  int eb_off; // offset of E_bath[bb] in the bath-energy array
};
struct AsmCon {
  int k, kp;
  long long fp_off;  // F_prev[p](ii_k, ii_kp): kept x kept, row-major, leading dim fp_ld
  int fp_ld;
  int fb_off;        // F_bath[p](bb_kp, bb_k): dbb_kp x dbb_k, row-major, leading dim fb_ld
  int fb_ld;
  double chi;        // chi_bath[bb_kp]
  double t;          // t_hopping[p]
};
struct AsmSector {
  long long h_off;  // element offset of the n x n block in the FP64 arena
  int n;
  int row0;         // first global row of this sector (prefix of n)
  int sub_begin, sub_end;
  int con_begin, con_end;
};

// Batched eigensolver blocks.
struct EigBlock {
  long long off;   // element offset of the n x n block (same in every per-block arena)
  int n;
  int nbc;         // number of block columns in the Jacobi tournament (even)
  int bw;          // block-column width
  int e_off;       // offset of the block's eigenvalues (prefix of n)
  int bc_off;      // offset of the block's block columns (prefix of nbc)
};

// One Jacobi task (a pair slot j of block b's tournament) with the block's fields, so that a
// workgroup reads everything it needs to start in one record.
struct JacTask {
  long long off;
  int n, nbc, bw, bc_off;
  int b, j;
};

// Batched GEMM problem: C = alpha * op(A) * op(B) + beta * C0, all row-major.
struct GemmProb {
  const double* A;
  const double* B;
  const double* C0;
  double* C;
  int m, n, k;
  int lda, ldb, ldc, ldc0;
  int transA, transB;
  double alpha, beta;
  int tile0;  // first tile index of this problem (prefix over problems)
  int tiles_n;
  int sym = 0;  // symmetric result: only the tiles on and above the diagonal are computed
};

// Operator rotation (k_rotate.hip) and truncation (k_trunc.hip) descriptors.
struct RotGroup {   // rows [kidx, kidx + kii*dbb) of sector i, sub-block (ii, bb), for one triple
  int triple;
  int kidx, kii, dbb;
  int seg_begin, seg_end;
};

struct RotSegment {  // a coupled sub-block (ii, bb') of sector j with F_bath[p](bb, bb') present
  int kidxp, dbbp;
  int fb_off, fb_ld;
};

struct RotTri {
  long long uj_off;  // U_j (n_j x n_j row-major) offset in the U arena
  int nj, kj;
  long long y_off;   // Y (n_i x kj row-major) offset in the scratch arena
};

struct ObsSeg {      // one nonzero sub-block of c(i,j)
  long long c_off;   // c(i,j) block (n_i x n_j row-major) offset
  int ldc;           // n_j
  int kidx, kidxp, kii, kiip, dbb;
  long long o_off;   // O_prev(ii, ii') block offset
  int ldo;           // its leading dimension (columns of the full previous block)
};

// FDM backward pass (k_fdm.hip)
struct FdmBinItem {
  long long b_off;    // B(i,j): n_i x n_j row-major
  long long ra_off;   // rhoA = B rho_j : n_i x kc_j (ld kc_j)
  long long ar_off;   // ARho = rho_i B : kc_i x n_j (ld n_j)
  int ni, nj, kci, kcj;
  int ei, ej;         // eigenvalue offsets of sectors i, j
  int op;
  int row0, rows;     // row tile
  int diag_mode;      // 1: rho diagonal with Boltzmann factors (last shell)
};
struct PtItem {       // reducedRho[ii] (kii x kii) and its contributions
  long long r_off;
  int kii;
  int c_begin, c_end;
};
struct PtContrib {    // sub-block (ii, bb) of sector i: rows kidx + it + kii*il
  long long rho_off;  // rho_i (n_i x n_i)
  int ni, kidx, dbb;
};
struct DiagItem {     // diagonal of block (i, i) of one system operator
  long long src;      // element offset of the n x n block, -1 if the operator has no such block
  long long dst;      // op * N_tot + first state of sector i
  int n;
};

// Wide-block eigensolver (k_tridiag.hip): blocked Householder tridiagonalisation in FP64, bisection
// for the eigenvalues, inverse iteration for the eigenvectors, WY back-transformation.  All n x n
// arenas of a block share the block's element offset a_off (row-major, leading dimension n):
//   A    working copy of H; the lower triangle is reduced in place; later the V T products
//   Vall Householder vectors v_k in column k (v_k[k+1] = 1, rows above zero)
//   U    the eigenvectors of T, then (in place) those of H, in columns 0 .. m-1
// The per-block workspace (w_off into the workspace arena) holds the panel's W (n x 32), the symv
// partials, the column under reduction and the tridiagonal (d, e, tau); see k_tridiag.hip.
constexpr int HT_TB = 64;  // tile edge of the symmetric matrix-vector product and the rank-2k update
constexpr int HT_NB = 32;  // panel width (reflectors per rank-2k update and per WY block)
struct HtBlock {
  long long a_off;  // element offset of the block's n x n arenas (H, U: row stride n)
  long long p_off;  // element offset of its working matrix and Householder vectors (row stride ld)
  long long w_off;  // element offset of its workspace
  long long s_off;  // element offset of its inverse-iteration scratch (3 n x mpad)
  int n;
  int T;            // HT_TB tiles per dimension
  int e_off;        // offset of the block's eigenvalues (prefix of n)
  int m;            // eigenvectors to compute (columns 0 .. m-1)
  int mpad;         // scratch slots of the inverse iteration for this block (ht_stein_slots(m))
  int ld;           // row stride of the working matrix: n rounded up to 16 (128-byte rows)
};

// The workspace of one block: W (n x HT_NB), the symv partials of the NT = T(T+1)/2 lower tiles
// (row sums rp and column sums cp, 64 each, and the v^T A v share sp), the per-tile-row partials of
// V^T v and W^T v (qv, qw: HT_NB each) and of the squared column norm (ss), the column under
// reduction (u) and the norm partials (ss) in two buffers each by column parity -- a launch reads
// column k's while it writes column k + 1's -- and the tridiagonal d, e with the reflector scales tau.
struct HtWs {
  double *W, *rp, *cp, *sp, *qv, *qw, *ss, *u, *d, *e, *tau;
};
__host__ __device__ inline long long ht_ws_size(int n, int T) {
  const long long NT = (long long)T * (T + 1) / 2;
  return (long long)n * HT_NB + NT * 129 + (long long)T * (2 * HT_NB + 2) + 5LL * n;
}
__host__ __device__ inline HtWs ht_ws(double* base, int n, int T) {
  const long long NT = (long long)T * (T + 1) / 2;
  HtWs w;
  w.W = base;
  w.rp = w.W + (long long)n * HT_NB;
  w.cp = w.rp + NT * 64;
  w.sp = w.cp + NT * 64;
  w.qv = w.sp + NT;
  w.qw = w.qv + (long long)T * HT_NB;
  w.ss = w.qw + (long long)T * HT_NB;
  w.u = w.ss + 2LL * T;
  w.d = w.u + 2LL * n;
  w.e = w.d + n;
  w.tau = w.e + n;
  return w;
}

struct TruncOut {
  double e_cut, e_min, e_max;
  int truncated;
  int n_kept;
  int near_ties;  // eigenvalues within the tie window of the cut e_cut + errbar (either side)
  int nonfinite;  // eigenvalues that are not finite (the step fails)
};

}  // namespace nrgk
