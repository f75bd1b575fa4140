// Chain orchestration: the device-resident counterpart of nrgcore<imp,bath>
// (nrgcore/include/nrgcore/nrgcore.hpp).  Integer bookkeeping (sectors, coupled index lists,
// descriptor tables) stays on the host; every floating-point stage runs on the chain's stream.
#include "chain.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>

#include "kernels.hpp"

namespace nrgk {

ThreadCounts& thread_counts() {
  static thread_local ThreadCounts tc;
  return tc;
}

// ------------------------------------------------------------------ DevBuf
void DevBuf::ensure(size_t bytes) {
  if (bytes <= cap) return;
  release();
  size_t want = bytes + bytes / 4 + 256;
  if (ctx) ctx->charge(want);
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    if (ctx) ctx->refund(want);
    p = nullptr;
    throw Error(NRGK_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(want) + "): " + hipGetErrorString(e));
  }
  cap = want;
}
void DevBuf::release() {
  if (p) {
    (void)hipFree(p);
    if (ctx) ctx->refund(cap);
  }
  p = nullptr;
  cap = 0;
}

template <class T>
static T* upload(DevBuf& buf, const std::vector<T>& v, hipStream_t s) {
  buf.ensure(std::max<size_t>(1, v.size()) * sizeof(T));
  h2d(buf.as<T>(), v.data(), v.size(), s);
  return buf.as<T>();
}

// ------------------------------------------------------------------ site models
SiteData SiteData::from_desc(const nrgk_site_desc* d) {
  NRGK_REQUIRE(d && d->n_sectors > 0 && d->n_qn > 0 && d->qn && d->dims && d->chi, NRGK_ERR_INVALID,
               "nrgk: site description is incomplete");
  SiteData s;
  s.S = d->n_sectors;
  s.nqn = d->n_qn;
  size_t eo = 0;
  for (int i = 0; i < s.S; i++) {
    s.qn.emplace_back(d->qn + (size_t)i * s.nqn, d->qn + (size_t)(i + 1) * s.nqn);
    NRGK_REQUIRE(d->dims[i] >= 0, NRGK_ERR_INVALID, "nrgk: negative sector dimension");
    s.E.emplace_back(d->energies + eo, d->energies + eo + d->dims[i]);
    eo += d->dims[i];
    s.chi.push_back(d->chi[i]);
  }
  s.nops = d->n_ops;
  s.ops.resize(s.nops);
  size_t bo = 0, dof = 0;
  for (int p = 0; p < s.nops; p++)
    for (int b = 0; b < d->op_nblocks[p]; b++, bo++) {
      const int bi = d->op_bi[bo], bj = d->op_bj[bo];
      NRGK_REQUIRE(bi >= 0 && bi < s.S && bj >= 0 && bj < s.S, NRGK_ERR_INVALID, "nrgk: operator block out of range");
      const size_t sz = (size_t)s.dim(bi) * s.dim(bj);
      NRGK_REQUIRE(!s.ops[p].count({bi, bj}), NRGK_ERR_INVALID, "Found operator already");
      s.ops[p][{bi, bj}] = std::vector<double>(d->op_data + dof, d->op_data + dof + sz);
      dof += sz;
    }
  return s;
}

// ------------------------------------------------------------------ chain
Chain::Chain(Ctx* c, const nrgk_site_desc* imp_d, const nrgk_site_desc* bath_d, size_t K_, int mode_)
    : ctx(c), K(K_), mode(mode_) {
  imp = SiteData::from_desc(imp_d);
  bath = SiteData::from_desc(bath_d);
  // nrgcore::test (nrgcore.hpp:117-134)
  NRGK_REQUIRE(imp.nqn == bath.nqn, NRGK_ERR_INVALID, "Number of quantum number is different on model and impurity!");
  NRGK_REQUIRE(imp.nops == bath.nops, NRGK_ERR_INVALID,
               "f_dag_operator.size is different on model and impurity!\nimpurityModel->f_dag_operator.size() = " +
                   std::to_string(imp.nops) + "bath_model->f_dag_operator.size() = " + std::to_string(bath.nops));
  init_device();
  upload_bath();
}

Chain::Chain(Ctx* c, int mode_) : ctx(c), mode(mode_) { init_device(); }

void Chain::init_device() {
  NRGK_REQUIRE(mode == NRGK_MODE_MIXED || mode == NRGK_MODE_PLAIN64, NRGK_ERR_INVALID, "nrgk: unknown mode");
  for (DevBuf* b : {&eprev_own, &fbuf[0], &fbuf[1], &d_bathE, &d_bathF, &dH, &dW, &dX, &dX2, &dP, &dS, &dR, &dE, &dU,
                    &dY, &dC, &dlam_tmp, &dlam_raw, &dlam_prev, &dlam_shift[0], &dlam_shift[1], &d_asm, &d_eig, &d_misc, &d_rot,
                    &d_gemm, &d_flags, &d_sigma_bits, &d_sigma, &d_stats, &d_kept, &d_trunc, &d_conv, &d_eig_act,
                    &dQ, &d_rrv, &d_cid, &d_sidx, &d_offmax, &d_dgram, &d_dgv, &d_ht, &d_wide_ws, &d_wide_s, &d_wide_q,
                    &d_wide_A, &d_wide_V})
    b->ctx = ctx;
  stg.ctx = ctx;
  NRGK_HIP(hipSetDevice(ctx->device));
  NRGK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  for (auto& e : ev) NRGK_HIP(hipEventCreate(&e));
}

// qmatrix::diag (qmatrix.hpp:375-404) for a list of blocks: the blocks are laid out as an
// iteration's sector blocks are (packed, offsets n^2 apart) and go through the same eigensolver.
void Chain::solve_blocks(const std::vector<int>& dims, const double* H, double* w, double* U) {
  prepare_blocks(dims, H);
  if (N_tot == 0) return;
  fail_on_unconverged = true;
  eigensolve();
  // a chain re-solves blocks that the refinement left short of its stopping point; the stand-alone
  // solver fails instead, as dsyevd reports a nonzero info
  NRGK_REQUIRE(mode == NRGK_MODE_PLAIN64 || oa_unconverged == 0, NRGK_ERR_NOCONV,
               "nrgk_eigh_batched: the FP64 refinement left " + std::to_string(oa_unconverged) +
                   " block(s) unconverged");
  NRGK_HIP(hipMemcpyAsync(w, dlam_raw.p, (size_t)N_tot * sizeof(double), hipMemcpyDeviceToHost, st));
  NRGK_HIP(hipMemcpyAsync(U, dU.p, (size_t)tot_elems * sizeof(double), hipMemcpyDeviceToHost, st));
  sync();
}

void Chain::prepare_blocks(const std::vector<int>& dims, const double* H) {
  const int S = (int)dims.size();
  cur_qn.assign(S, std::vector<int>(1, 0));
  cur_n = dims;
  cur_eoff.assign(S, 0);
  cur_off.assign(S, 0);
  tot_elems = 0;
  N_tot = 0;
  nmax = 0;
  for (int s = 0; s < S; s++) {
    NRGK_REQUIRE(dims[s] >= 0, NRGK_ERR_INVALID, "nrgk_eigh_batched: negative block size");
    cur_eoff[s] = (int)N_tot;
    cur_off[s] = tot_elems;
    N_tot += dims[s];
    tot_elems += (long long)dims[s] * dims[s];
    nmax = std::max(nmax, dims[s]);
  }
  if (N_tot == 0) return;
  NRGK_HIP(hipSetDevice(ctx->device));
  // LAPACKE_dsyevd(ROW_MAJOR, 'U') reads only the upper triangle of each block (qmatrix.hpp:390-391):
  // it is mirrored into the lower one, which the batched solver also reads
  std::vector<double> Hs(H, H + tot_elems);
  for (int s = 0; s < S; s++) {
    double* A = Hs.data() + cur_off[s];
    const int n = dims[s];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < i; j++) A[(size_t)i * n + j] = A[(size_t)j * n + i];
  }
  dH.ensure((size_t)tot_elems * sizeof(double));
  NRGK_HIP(hipMemcpyAsync(dH.p, Hs.data(), (size_t)tot_elems * sizeof(double), hipMemcpyHostToDevice, st));
  // the host copy must outlive the asynchronous copy
  NRGK_HIP(stream_sync(st));
}

void Chain::sync() {
  NRGK_HIP(stream_sync(st));
  stg.reset();
  prof_resolve();
}

hipEvent_t Chain::take_event() {
  if (ev_pool.empty()) {
    hipEvent_t e;
    NRGK_HIP(hipEventCreate(&e));
    return e;
  }
  hipEvent_t e = ev_pool.back();
  ev_pool.pop_back();
  return e;
}
void Chain::prof_begin(const char* family, hipStream_t s) {
  if (!prof_on || (!prof_only.empty() && prof_only != family)) return;
  prof_open = take_event();
  NRGK_HIP(hipEventRecord(prof_open, s ? s : st));
}
void Chain::prof_end(const char* family, long long launches, double flops, double bytes, hipStream_t s) {
  if (!prof_on || !prof_open) return;
  hipEvent_t e1 = take_event();
  NRGK_HIP(hipEventRecord(e1, s ? s : st));
  prof_pending.push_back(ProfRun{family, prof_open, e1, launches, flops, bytes});
  prof_open = nullptr;
}
void Chain::prof_resolve() {
  for (auto& r : prof_pending) {
    float t = 0;
    NRGK_HIP(hipEventSynchronize(r.e1));
    NRGK_HIP(hipEventElapsedTime(&t, r.e0, r.e1));
    KStat& k = kstats[r.family];
    k.ms += t;
    k.launches += r.launches;
    k.flops += r.flops;
    k.bytes += r.bytes;
    ev_pool.push_back(r.e0);
    ev_pool.push_back(r.e1);
  }
  prof_pending.clear();
}

Chain::~Chain() {
  wide_thread_join();
  if (st) (void)hipStreamSynchronize(st);
  plain_destroy(*this);
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (auto& r : prof_pending) {
    ev_pool.push_back(r.e0);
    ev_pool.push_back(r.e1);
  }
  if (prof_open) ev_pool.push_back(prof_open);
  for (auto& e : ev_pool) (void)hipEventDestroy(e);
  if (st) (void)hipStreamDestroy(st);
}

void Chain::set_bath(const nrgk_site_desc* b) {
  SiteData nb = SiteData::from_desc(b);
  NRGK_REQUIRE(nb.nqn == imp.nqn && nb.nops == imp.nops, NRGK_ERR_INVALID, "nrgk: bath model does not match the impurity");
  bath = std::move(nb);
  upload_bath();
}

void Chain::upload_bath() {
  std::vector<double> e;
  bath_eoff.assign(bath.S, 0);
  for (int s = 0; s < bath.S; s++) {
    bath_eoff[s] = (int)e.size();
    e.insert(e.end(), bath.E[s].begin(), bath.E[s].end());
  }
  upload(d_bathE, e, st);
  std::vector<double> f;
  bathF.assign(bath.nops, OpMap());
  for (int p = 0; p < bath.nops; p++)
    for (auto& [key, blk] : bath.ops[p]) {
      bathF[p][key] = BlockRef{(long long)f.size(), bath.dim(key.first), bath.dim(key.second)};
      f.insert(f.end(), blk.begin(), blk.end());
    }
  upload(d_bathF, f, st);
  sync();
}

// First add_bath_site: impurity data and discard_higher_energies on it (nrgcore.hpp:72-78).
void Chain::import_impurity() {
  pre_qn = imp.qn;
  std::vector<double> allE;
  for (auto& v : imp.E) allE.insert(allE.end(), v.begin(), v.end());
  NRGK_REQUIRE(!allE.empty(), NRGK_ERR_INVALID, "nrgk: empty impurity");
  std::vector<double> sorted = allE;
  std::sort(sorted.begin(), sorted.end());
  all_eig = sorted;
  const double emin = sorted[0];
  pre_kept.assign(imp.S, 0);
  std::vector<std::vector<double>> E = imp.E;
  if (sorted.size() <= K) {
    for (int s = 0; s < imp.S; s++) pre_kept[s] = imp.dim(s);
  } else {
    const double emax = sorted[K];
    for (int s = 0; s < imp.S; s++) {
      int k = 0;
      for (double x : imp.E[s])
        if (x < emax + errbar) k++;
      pre_kept[s] = k;
    }
  }
  no_kept = 0;
  for (int k : pre_kept) no_kept += k;
  if (no_kept <= K) {
    cnt_min = cnt;
  } else {
    relgs.push_back(emin);
    for (auto& v : E)
      for (auto& x : v) x = x - emin;
  }
  std::vector<double> flat;
  pre_eoff.assign(imp.S, 0);
  for (int s = 0; s < imp.S; s++) {
    pre_eoff[s] = (int)flat.size();
    flat.insert(flat.end(), E[s].begin(), E[s].end());
  }
  d_eprev = upload(eprev_own, flat, st);
  std::vector<double> f;
  fprev.assign(imp.nops, OpMap());
  for (int p = 0; p < imp.nops; p++)
    for (auto& [key, blk] : imp.ops[p]) {
      fprev[p][key] = BlockRef{(long long)f.size(), imp.dim(key.first), imp.dim(key.second)};
      f.insert(f.end(), blk.begin(), blk.end());
    }
  upload(fbuf[fcur], f, st);
  h_last_shift = flat;
}

// nrgcore::create_next_basis (nrgcore.hpp:238-261)
void Chain::create_next_basis() {
  cur_qn.clear();
  coupled.clear();
  // sectors in first-seen order over (i, j); a hash of the quantum numbers replaces the reference's
  // linear search.  Up to four quantum numbers in [-32768, 32767] pack into one 64-bit key (open
  // addressing, linear probing); other labels go through an ordered map.
  const size_t combos = pre_qn.size() * (size_t)bath.S;
  bool packable = bath.nqn <= 4;
  for (int x = 0; x < bath.nqn && packable; x++) {
    long long lo = 0, hi = 0, blo = 0, bhi = 0;
    for (auto& q : pre_qn) lo = std::min<long long>(lo, q[x]), hi = std::max<long long>(hi, q[x]);
    for (auto& q : bath.qn) blo = std::min<long long>(blo, q[x]), bhi = std::max<long long>(bhi, q[x]);
    packable = lo + blo >= -32768 && hi + bhi <= 32767;
  }
  auto pack = [&](const std::vector<int>& q) {
    unsigned long long k = 0;
    for (int x = 0; x < bath.nqn; x++) k = (k << 16) | (unsigned long long)(unsigned short)(q[x] + 32768);
    return k;
  };
  size_t cap = 16;
  while (cap < 2 * combos) cap <<= 1;
  std::vector<unsigned long long> hkey(cap);
  std::vector<int> hval(cap, -1);
  std::map<std::vector<int>, int> index;
  std::vector<int> tm;
  for (size_t i = 0; i < pre_qn.size(); i++)
    for (int j = 0; j < bath.S; j++) {
      tm = pre_qn[i];
      for (int x = 0; x < bath.nqn; x++) tm[x] += bath.qn[j][x];
      const unsigned long long key = packable ? pack(tm) : 0;
      int found = -1;
      size_t slot = 0;
      if (packable) {
        slot = (size_t)((key * 0x9E3779B97F4A7C15ull) >> 20) & (cap - 1);
        while (hval[slot] >= 0 && hkey[slot] != key) slot = (slot + 1) & (cap - 1);
        found = hval[slot];
      } else {
        auto it = index.find(tm);
        if (it != index.end()) found = it->second;
      }
      if (found >= 0) {
        coupled[found].push_back({(int)i, j});
      } else {
        if (packable) {
          hkey[slot] = key;
          hval[slot] = (int)cur_qn.size();
        } else {
          index.emplace(tm, (int)cur_qn.size());
        }
        cur_qn.push_back(tm);
        coupled.push_back({{(int)i, j}});
      }
    }
  const int S = (int)cur_qn.size();
  cur_n.assign(S, 0);
  cur_eoff.assign(S, 0);
  cur_off.assign(S, 0);
  sub_kidx.assign(S, {});
  tot_elems = 0;
  N_tot = 0;
  nmax = 0;
  for (int s = 0; s < S; s++) {
    int n = 0;
    for (auto [ii, bb] : coupled[s]) {
      sub_kidx[s].push_back(n);
      n += pre_kept[ii] * bath.dim(bb);
    }
    cur_n[s] = n;
    cur_eoff[s] = (int)N_tot;
    cur_off[s] = tot_elems;
    N_tot += n;
    tot_elems += (long long)n * n;
    nmax = std::max(nmax, n);
  }
}

// K1 launch (nrgcore::create_next_hamiltonians, nrgcore.hpp:146-230)
void Chain::assemble(const std::vector<double>& t, double rescale) {
  const int S = (int)cur_qn.size();
  std::vector<AsmSector> secs;
  std::vector<AsmSub> subs;
  std::vector<AsmCon> cons;
  int row0 = 0;
  for (int s = 0; s < S; s++) {
    if (cur_n[s] == 0) continue;
    AsmSector sec{};
    sec.h_off = cur_off[s];
    sec.n = cur_n[s];
    sec.row0 = row0;
    row0 += cur_n[s];
    sec.sub_begin = (int)subs.size();
    const auto& cl = coupled[s];
    for (size_t k = 0; k < cl.size(); k++) {
      auto [ii, bb] = cl[k];
      subs.push_back(AsmSub{sub_kidx[s][k], pre_kept[ii], bath.dim(bb), pre_eoff[ii], bath_eoff[bb]});
    }
    sec.sub_end = (int)subs.size();
    sec.con_begin = (int)cons.size();
    for (size_t k = 0; k < cl.size(); k++) {
      auto [ii, bb] = cl[k];
      if (pre_kept[ii] == 0 || bath.dim(bb) == 0) continue;
      for (size_t kp = 0; kp < cl.size(); kp++) {
        auto [iip, bbp] = cl[kp];
        if (pre_kept[iip] == 0 || bath.dim(bbp) == 0) continue;
        for (int p = 0; p < imp.nops; p++) {
          auto fp = fprev[p].find({ii, iip});
          auto fb = bathF[p].find({bbp, bb});
          if (fp == fprev[p].end() || fb == bathF[p].end()) continue;
          AsmCon c{};
          c.k = (int)k;
          c.kp = (int)kp;
          c.fp_off = fp->second.off;
          c.fp_ld = fp->second.cols;
          c.fb_off = (int)fb->second.off;
          c.fb_ld = fb->second.cols;
          c.chi = bath.chi[bbp];
          c.t = t[p];
          cons.push_back(c);
        }
      }
    }
    sec.con_end = (int)cons.size();
    secs.push_back(sec);
  }
  Staging::Region rg = stg.open(Staging::span<AsmSector>(secs.size()) + Staging::span<AsmSub>(subs.size()) +
                                Staging::span<AsmCon>(cons.size()));
  const AsmSector* d_secs = stg.place(rg, secs);
  const AsmSub* d_subs = stg.place(rg, subs);
  const AsmCon* d_cons = stg.place(rg, cons);
  stg.commit(rg, st);
  dH.ensure(std::max<long long>(1, tot_elems) * sizeof(double));
  // algorithmic bytes: every element of the n x n blocks written once (8 B), every kept x kept
  // f-dagger block and bath block a connection reads once
  double by = 0;
  if (prof_on) {
    by = 8.0 * (double)tot_elems;
    for (const auto& sec : secs)
      for (int k = sec.con_begin; k < sec.con_end; k++) {
        const AsmCon& c = cons[k];
        const AsmSub& sk = subs[sec.sub_begin + c.k];
        const AsmSub& skp = subs[sec.sub_begin + c.kp];
        by += 8.0 * ((double)sk.kii * skp.kii + (double)sk.dbb * skp.dbb);
      }
  }
  prof_begin("assemble");
  launch_assemble(d_secs, (int)secs.size(), row0, d_subs, d_cons, d_eprev, d_bathE.as<double>(),
                  fbuf[fcur].as<double>(), d_bathF.as<double>(), rescale, dH.as<double>(), st);
  prof_end("assemble", 1, 0.0, by);
}

void Chain::run_gemms(std::vector<GemmProb>& probs) {
  if (mode == NRGK_MODE_PLAIN64) {
    plain_gemms(*this, probs);
    return;
  }
  // products of blocks of 1024 and wider are plain library GEMMs (rocBLAS dgemm, one call each, on
  // the chain's stream); the many smaller ones share one launch of the batched kernel
  std::vector<GemmProb> live, large;
  for (auto& p : probs)
    if (p.m > 0 && p.n > 0) (p.m >= 1024 && p.n >= 1024 && p.k >= 1024 ? large : live).push_back(p);
  if (!large.empty()) {
    double fl = 0;
    for (auto& p : large) fl += 2.0 * p.m * p.n * p.k;
    prof_begin("gemm_rocblas");
    plain_gemms(*this, large);  // synchronises the stream
    prof_end("gemm_rocblas", (long long)large.size(), fl, 0.0);
  }
  if (live.empty()) return;
  const int tiles = prepare_gemm(live);
  const GemmProb* dprobs = stg.put(live, st);
  double fl = 0, by = 0;
  if (prof_on)
    for (auto& p : live) {
      // a symmetric product computes (and counts) the n (n + 1) / 2 entries on and above the diagonal
      const double outs = p.sym ? 0.5 * (double)p.m * (p.m + 1) : (double)p.m * p.n;
      fl += 2.0 * outs * p.k;
      by += 8.0 * ((double)p.m * p.k + (double)p.k * p.n + (p.C0 ? 2.0 : 1.0) * outs);
    }
  prof_begin("gemm_f64");
  launch_gemm(dprobs, (int)live.size(), tiles, st);
  prof_end("gemm_f64", 1, fl, by);
}

// tuning knobs read from the environment (defaults are the tuned values)
static double knob(const char* name, double dflt) {
  const char* v = std::getenv(name);
  return v ? std::atof(v) : dflt;
}

// block-column width of the Jacobi tournament: 16 (a task's Gram matrix is 32 x 32), or half the
// block for blocks of at most 32
static int choose_bw(int n) { return n <= 32 ? std::max(1, (n + 1) / 2) : 16; }

// K2 + K6: FP32 block Jacobi, then two Ogita-Aishima refinement steps in FP64.
void Chain::eigensolve_mixed() {
  const int S = (int)cur_qn.size();
  // blocks wider than the crossover (the context's, tuned or 768; NRGK_MIXED_MAX overrides it) go to
  // the wide-block path on the side stream
  const int mixed_max = mixed_max_override >= 0 ? mixed_max_override : (int)knob("NRGK_MIXED_MAX", ctx->mixed_max);
  std::vector<int> big;
  for (int s = 0; s < S; s++)
    if (cur_n[s] > mixed_max) big.push_back(s);
  const size_t bytes64_all = std::max<long long>(1, tot_elems) * sizeof(double);
  dX.ensure(bytes64_all);  // every buffer the side stream touches is sized before it starts
  dX2.ensure(bytes64_all);
  dU.ensure(bytes64_all);
  dlam_raw.ensure(std::max<long long>(1, N_tot) * sizeof(double));
  // the wide blocks: the FP64 Householder path (default) or, with NRGK_WIDE_SOLVER=rocsolver,
  // rocSOLVER dsyevd per block on the side streams
  const char* wsv = std::getenv("NRGK_WIDE_SOLVER");
  bool wide_rocsolver = wsv && std::string(wsv) == "rocsolver";
  // blocks wider than the tridiagonal kernels' LDS (over 9 000 rows) take rocSOLVER
  for (int s : big) wide_rocsolver = wide_rocsolver || cur_n[s] > ht_max_n(ctx->lds_max);
  ht.clear();
  wide_sectors.clear();
  wide_state = WIDE_NONE;
  if (wide_rocsolver) {
    large_blocks_begin(*this, big);
  } else {
    large_blocks_begin(*this, {});
    wide_begin(big);
  }
  std::vector<EigBlock> blks;
  std::vector<int> row_block, row_local;
  std::vector<int2> tasks;
  int nbc_max = 2, n_bc = 0;
  for (int s = 0; s < S; s++) {
    const int n = cur_n[s];
    if (n == 0 || n > mixed_max) continue;
    EigBlock b{};
    b.off = cur_off[s];
    b.n = n;
    b.bw = choose_bw(n);
    int nbc = (n + b.bw - 1) / b.bw;
    if (nbc < 2) nbc = 2;
    if (nbc & 1) nbc++;
    b.nbc = nbc;
    b.e_off = cur_eoff[s];
    b.bc_off = n_bc;
    n_bc += nbc;
    const int bi = (int)blks.size();
    for (int j = 0; j < nbc / 2; j++) tasks.push_back(make_int2(bi, j));
    for (int r = 0; r < n; r++) {
      row_block.push_back(bi);
      row_local.push_back(r);
    }
    nbc_max = std::max(nbc_max, nbc);
    blks.push_back(b);
  }
  const int nb = (int)blks.size();
  if (nb == 0) {  // every block went to the side streams: no batched launch (empty grids)
    large_blocks_join(*this);
    wide_join();
    sync();
    large_blocks_check(*this);
    sweeps = maxcluster = unresolved = oa_steps = oa_unconverged = 0;
    if (need_full_u || fail_on_unconverged) {
      wide_vectors(true);
      sync();
    }
    return;
  }
  const int nrows = (int)row_block.size();
  int nmax_m = 1;  // widest block of the batched solve (LDS sizes)
  for (auto& b : blks) nmax_m = std::max(nmax_m, b.n);
  upload_to(d_eig, blks.data(), nb);
  const size_t o_rb = 0, o_rl = ((nrows * 4 + 255) / 256) * 256;
  d_misc.ensure(o_rl + std::max(1, nrows) * sizeof(int));
  upload_to(d_misc, row_block.data(), nrows, o_rb);
  upload_to(d_misc, row_local.data(), nrows, o_rl);
  char* mb = d_misc.as<char>();
  const EigBlock* dblk = d_eig.as<EigBlock>();

  const size_t bytes64 = std::max<long long>(1, tot_elems) * sizeof(double);
  dW.ensure(std::max<long long>(1, tot_elems) * sizeof(float));
  d_sigma_bits.ensure(std::max(1, nb) * 2 * sizeof(unsigned));
  d_sigma.ensure(std::max(1, nb) * sizeof(float));
  NRGK_HIP(hipMemsetAsync(d_sigma_bits.p, 0, std::max(1, nb) * 2 * sizeof(unsigned), st));
  launch_shift_convert(dblk, (const int*)(mb + o_rb), (const int*)(mb + o_rl), nrows, dH.as<double>(),
                       d_sigma_bits.as<unsigned>(), d_sigma.as<float>(), dW.as<float>(), st);
  // Jacobi sweeps.  The tasks of the blocks still rotating are listed per sweep (a block that a sweep
  // leaves unrotated has converged), blocks with the longest tournaments first, so that round r
  // launches the prefix of tasks whose block has a round r.
  const int max_sweeps = 20;
  // Gram blocks of the block columns, carried between rounds (invalid until first formed)
  d_dgram.ensure(std::max(1, n_bc) * 256 * sizeof(double));
  d_dgv.ensure(std::max(1, n_bc) * sizeof(int));
  NRGK_HIP(hipMemsetAsync(d_dgv.p, 0, std::max(1, n_bc) * sizeof(int), st));
  // rotate a column pair while cos > jac_tol and the rotation angle tan > jac_ang: the angle test
  // stops pairs of very different norms at the FP32 floor, the cosine test degenerate pairs
  const float jac_tol = knob("NRGK_JAC_TOL", 5e-7);
  const float jac_ang = knob("NRGK_JAC_ANGLE", 1e-7);
  const int jac_inner = (int)knob("NRGK_JAC_INNER", 1);  // Jacobi sweeps over a task's Gram matrix
  // stop a block after a sweep whose pairs all started near orthogonal (0: only after a sweep
  // without rotations)
  const bool jac_quad = knob("NRGK_JAC_QUAD", 1) != 0;
  const double jac_far = knob("NRGK_JAC_FAR", 1e-4);
  const bool dbg_jac = knob("NRGK_DEBUG_JAC", 0) != 0;
  // Workgroup g of a launch runs on XCD g mod 8 (round-robin dispatch).  Every block is kept on one
  // XCD for the whole solve, so that its columns stay in that XCD's L2 from round to round: blocks,
  // longest tournaments first, go to the XCD with the fewest tasks so far, and slot k of XCD x's
  // task list sits at grid index 8 k + x (unused slots hold an empty task, which returns at once).
  constexpr int NXCD = 8;
  std::vector<int> order(nb);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return blks[x].nbc > blks[y].nbc; });
  // The balance counts the tasks that do work: after round 0 the pair holding the empty block column
  // that pads an odd count returns at once.  Blocks with rounds after the first are balanced on their
  // tasks per later round (32 CUs per XCD: more than 32 tasks on one XCD need a second wave), the
  // one-round blocks on their round-0 tasks.
  std::vector<int> xcd_of(std::max(1, nb), 0);
  {
    int load[NXCD] = {0}, load0[NXCD] = {0};
    for (int bi : order) {
      const EigBlock& e = blks[bi];
      const int w = e.nbc > 2 ? e.nbc / 2 - ((e.nbc - 1) * e.bw >= e.n ? 1 : 0) : 0;
      int x = 0;
      for (int y = 1; y < NXCD; y++) {
        const bool better = w > 0 ? (load[y] < load[x] || (load[y] == load[x] && load0[y] < load0[x]))
                                  : (load0[y] < load0[x] || (load0[y] == load0[x] && load[y] < load[x]));
        if (better) x = y;
      }
      xcd_of[bi] = x;
      load[x] += w;
      load0[x] += e.nbc / 2;
    }
  }
  std::vector<int> hflags(std::max(1, nb), 1);
  std::vector<JacTask> xtasks[NXCD], grid_tasks;
  sweeps = 0;
  bool converged = false;
  for (int sw = 0; sw < max_sweeps && nb > 0; sw++) {
    for (auto& v : xtasks) v.clear();
    int nbc_act = 2;
    double fl = 0, by = 0;
    for (int bi : order) {
      if (!hflags[bi]) continue;
      const EigBlock& e = blks[bi];
      nbc_act = std::max(nbc_act, e.nbc);
      for (int jj = 0; jj < e.nbc / 2; jj++)
        xtasks[xcd_of[bi]].push_back(JacTask{e.off, e.n, e.nbc, e.bw, e.bc_off, bi, jj});
      if (prof_on) {  // per task and round: the cross Gram block (2 n (w/2)^2) and P V (2 n w^2)
        const double n = e.n, w = std::min(e.n, 2 * e.bw);
        const double rounds = e.nbc - 1, pairs = std::ceil(n / w);
        fl += rounds * pairs * 2.5 * n * w * w;
        by += rounds * pairs * (n <= 768 ? 2.0 : 3.0) * n * w * 4.0;  // FP32 panel read once, written once
      }
    }
    size_t maxlen = 0;
    for (auto& v : xtasks) maxlen = std::max(maxlen, v.size());
    grid_tasks.assign(NXCD * maxlen, JacTask{0, 0, 0, 1, 0, 0, 0});
    for (int x = 0; x < NXCD; x++)
      for (size_t k = 0; k < xtasks[x].size(); k++) grid_tasks[NXCD * k + x] = xtasks[x][k];
    Staging::Region rg = stg.open(Staging::span<JacTask>(grid_tasks.size()) + Staging::span<int>(2 * nb));
    const JacTask* dt = stg.place(rg, grid_tasks);
    int* dflags = stg.place_zero<int>(rg, 2 * nb);
    stg.commit(rg, st);
    prof_begin("jacobi_round");
    size_t cnt[NXCD];
    for (int x = 0; x < NXCD; x++) cnt[x] = xtasks[x].size();
    for (int r = 0; r < nbc_act - 1; r++) {
      size_t len = 0;
      for (int x = 0; x < NXCD; x++) {  // tasks of blocks whose tournament has ended drop off the end
        while (cnt[x] > 0 && xtasks[x][cnt[x] - 1].nbc - 1 <= r) cnt[x]--;
        len = std::max(len, cnt[x]);
      }
      launch_jacobi_round(dt, (int)(NXCD * len), dW.as<float>(), r, dflags, d_dgram.as<double>(), d_dgv.as<int>(),
                          jac_tol, jac_ang, jac_inner, jac_far, st);
    }
    prof_end("jacobi_round", nbc_act - 1, fl, by);
    const int* hf = stg.get(dflags, 2 * nb, st);
    sync();
    sweeps = sw + 1;
    // a block goes on while it rotated and some pair was still far from orthogonal
    bool any = false;
    for (int bi = 0; bi < nb; bi++) {
      hflags[bi] = hflags[bi] && hf[2 * bi] && (jac_quad ? hf[2 * bi + 1] : 1);
      any |= hflags[bi] != 0;
    }
    if (dbg_jac) {  // NRGK_DEBUG_JAC=1: blocks still rotating (and their largest width) after each sweep
      int act = 0, wmax = 0;
      for (int bi = 0; bi < nb; bi++)
        if (hflags[bi]) act++, wmax = std::max(wmax, blks[bi].n);
      std::fprintf(stderr, "nrgk jacobi sweep %d: %d of %d blocks go on (widest %d)\n", sw, act, nb, wmax);
    }
    if (!any) {
      converged = true;
      break;
    }
  }
  NRGK_REQUIRE(converged || nb == 0, NRGK_ERR_NOCONV, "nrgk: FP32 Jacobi did not converge in 20 sweeps");
  dX.ensure(bytes64);
  dX2.ensure(bytes64);
  dP.ensure(bytes64);
  dS.ensure(bytes64);
  dR.ensure(bytes64);
  dE.ensure(bytes64);
  dU.ensure(bytes64);
  dQ.ensure(bytes64);
  d_rrv.ensure(std::max<long long>(1, N_tot) * 3 * sizeof(double));
  dlam_tmp.ensure(std::max<long long>(1, N_tot) * sizeof(double));
  dlam_raw.ensure(std::max<long long>(1, N_tot) * sizeof(double));
  // 32 x 32 tiles of every block, for the gathers into X and U, then the 32-column groups for the
  // column norms
  std::vector<int4> tiles_blk;
  for (int bi = 0; bi < nb; bi++) {
    const int T = (blks[bi].n + 31) / 32;
    for (int ti = 0; ti < T; ti++)
      for (int tj = 0; tj < T; tj++) tiles_blk.push_back(int4{bi, ti, tj, 0});
  }
  const size_t n_tiles_blk = tiles_blk.size();
  for (int bi = 0; bi < nb; bi++)
    for (int c0 = 0; c0 < blks[bi].n; c0 += 32) tiles_blk.push_back(int4{bi, 0, c0, 0});
  const int4* d_tiles_blk = upload_to(d_eig_act, tiles_blk.data(), tiles_blk.size());
  const int4* d_colg = d_tiles_blk + n_tiles_blk;
  const int n_colg = (int)(tiles_blk.size() - n_tiles_blk);
  d_cid.ensure(std::max<long long>(1, N_tot) * sizeof(int));
  d_sidx.ensure(std::max<long long>(1, N_tot) * sizeof(int));
  launch_normalize_sort(dblk, nb, d_tiles_blk, (int)n_tiles_blk, d_colg, n_colg, dW.as<float>(), d_sigma.as<float>(),
                        dX.as<double>(), dlam_tmp.as<double>(), d_cid.as<int>(), d_rrv.as<double>(),
                        d_rrv.as<double>() + std::max<long long>(1, N_tot), normalize_sort_lds(nmax_m), st);
  // NRGK_DEBUG_FINITE=1: the first stage and block that produces a non-finite value fail the call
  // with NRGK_ERR_NOCONV (the stage and block are in the message)
  const bool dbg = knob("NRGK_DEBUG_FINITE", 0) != 0;
  auto check = [&](const char* stage, const DevBuf& buf, bool is_float, bool per_block_vec) {
    if (!dbg) return;
    NRGK_HIP(stream_sync(st));
    for (auto& b : blks) {
      const size_t cnt = per_block_vec ? (size_t)b.n : (size_t)b.n * b.n;
      const size_t off = per_block_vec ? (size_t)b.e_off : (size_t)b.off;
      std::vector<double> v(cnt);
      if (is_float) {
        std::vector<float> f(cnt);
        d2h(f.data(), buf.as<float>() + off, cnt, st);
        NRGK_HIP(stream_sync(st));
        for (size_t i = 0; i < cnt; i++) v[i] = f[i];
      } else {
        d2h(v.data(), buf.as<double>() + off, cnt, st);
        NRGK_HIP(stream_sync(st));
      }
      for (double x : v)
        if (!std::isfinite(x))
          throw Error(NRGK_ERR_NOCONV, std::string("nrgk: non-finite value after ") + stage + " in block n=" +
                                           std::to_string(b.n) + " off=" + std::to_string(b.off));
    }
  };
  check("shift_convert+jacobi (W)", dW, true, false);
  check("normalize_sort (X)", dX, false, false);
  check("normalize_sort (lam)", dlam_tmp, false, true);
  d_stats.ensure(4 * sizeof(int));
  NRGK_HIP(hipMemsetAsync(d_stats.p, 0, 4 * sizeof(int), st));
  // Ogita-Aishima steps, per block until the out-of-cluster rotation residuals |rho_ij| are at the
  // rounding noise of S (sqrt(n) eps max|lt|: no further step can lower them) and the orthogonality
  // defect is below 1e-12.  The eigenvectors then carry errors of noise/gap, as a backward-stable
  // FP64 solver's do; the next iteration's Hamiltonian is built from them, so a looser stop would
  // show up there.  Converged blocks leave the batch; both X buffers then hold their final vectors.
  const int oa_max = (int)knob("NRGK_OA_ITERS", 12);
  const double e_tol = knob("NRGK_OA_ETOL", 1.0), d_tol = knob("NRGK_OA_DTOL", 1e-12);
  // a full step from residuals within e_pred noise and a defect within d_pred is the block's final
  // one (its update lands at the noise; the verifying step is skipped)
  const double e_pred = knob("NRGK_OA_EPRED", 300.0), d_pred = knob("NRGK_OA_DPRED", 1e-6);
  const int dbg_oa = (int)knob("NRGK_DEBUG_OA", 0);
  std::vector<EigBlock> act = blks;
  oa_steps = 0;
  for (int it = 0; it < oa_max && !act.empty(); it++) {
    const int na = (int)act.size();
    // 32 x 32 tiles of the active blocks: all of them, and the (i <= j) half for the symmetrisation
    std::vector<int4> tiles_all, tiles_sym;
    for (int a = 0; a < na; a++) {
      const int T = (act[a].n + 31) / 32;
      for (int ti = 0; ti < T; ti++)
        for (int tj = 0; tj < T; tj++) {
          tiles_all.push_back(int4{a, ti, tj, 0});
          if (ti <= tj) tiles_sym.push_back(int4{a, ti, tj, 0});
        }
    }
    Staging::Region rg = stg.open(Staging::span<EigBlock>(act.size()) + Staging::span<int4>(tiles_all.size()) +
                                  Staging::span<int4>(tiles_sym.size()) + Staging::span<double>(4 * na) +
                                  Staging::span<double>(na));
    const EigBlock* m_blk = stg.place(rg, act);
    const int4* m_all = stg.place(rg, tiles_all);
    const int4* m_sym = stg.place(rg, tiles_sym);
    double* m_conv = stg.place_zero<double>(rg, 4 * na);
    double* m_offmax = stg.place_zero<double>(rg, na);
    stg.commit(rg, st);
    std::vector<GemmProb> g1, g2, g3;
    for (auto& b : act) {
      const int n = b.n;
      double* H = dH.as<double>() + b.off;
      double* X = dX.as<double>() + b.off;
      double* P = dP.as<double>() + b.off;
      double* Sm = dS.as<double>() + b.off;
      double* R = dR.as<double>() + b.off;
      double* E = dE.as<double>() + b.off;
      double* X2 = dX2.as<double>() + b.off;
      g1.push_back(GemmProb{H, X, nullptr, P, n, n, n, n, n, n, n, 0, 0, 1.0, 0.0, 0, 0});
      g1.push_back(GemmProb{X, X, nullptr, R, n, n, n, n, n, n, n, 1, 0, 1.0, 0.0, 0, 0, 1});
      g2.push_back(GemmProb{X, P, nullptr, Sm, n, n, n, n, n, n, n, 1, 0, 1.0, 0.0, 0, 0, 1});
      g3.push_back(GemmProb{X, E, X, X2, n, n, n, n, n, n, n, 0, 0, 1.0, 1.0, 0, 0});
    }
    run_gemms(g1);
    run_gemms(g2);
    check("P=HX", dP, false, false);
    check("S=X^T P", dS, false, false);
    check("R=X^T X", dR, false, false);
    // clusters wider than the LDS Rayleigh-Ritz work in P (free after S = X^T P), X2 (written only
    // by the X + X E product that follows) and Q
    double oa_bytes = 0;  // symmetrise S, Rp (read and write both), E from S and Rp, masking pass
    if (prof_on)
      for (auto& b : act) oa_bytes += 8.0 * 9.0 * (double)b.n * b.n;
    prof_begin("oa_update");
    launch_oa(m_blk, na, m_sym, (int)tiles_sym.size(), m_all, (int)tiles_all.size(), dS.as<double>(), dR.as<double>(),
              dE.as<double>(), dlam_tmp.as<double>(), d_cid.as<int>(), d_sidx.as<int>(), m_offmax,
              d_stats.as<int>(), m_conv, dP.as<double>(), dX2.as<double>(), dQ.as<double>(),
              d_rrv.as<double>(), nmax_m, e_tol, d_tol, st);
    prof_end("oa_update", 5, 0.0, oa_bytes);
    check("oa_update (E)", dE, false, false);
    check("oa_update (lam)", dlam_tmp, false, true);
    const double* h_conv = stg.get(m_conv, 4 * na, st);
    run_gemms(g3);
    sync();
    const std::vector<double> conv(h_conv, h_conv + 4 * na);
    std::swap(dX, dX2);
    check("X+XE", dX, false, false);
    oa_steps = it + 1;
    std::vector<EigBlock> next;
    for (int a = 0; a < na; a++) {
      // NRGK_DEBUG_OA=n: per step, the residual (in noise units) and orthogonality defect of every
      // active block at least n wide
      if (dbg_oa > 0 && act[a].n >= dbg_oa)
        std::fprintf(stderr, "nrgk oa step %d block off=%lld n=%d residual/noise=%.3g orth=%.3g rotation=%.3g\n", it,
                     act[a].off, act[a].n, conv[4 * a] / conv[4 * a + 2], conv[4 * a + 1], conv[4 * a + 3]);
      // the device's test (k_copy_converged) on the same numbers: the last step, or a full step whose
      // update is predicted to land at the noise (out-of-cluster rotations r with n r^2 <= d_tol, so
      // that X + XE loses no orthogonality beyond d_tol at second order)
      const double rot = conv[4 * a + 3];
      const bool last = conv[4 * a] <= e_tol * conv[4 * a + 2] && conv[4 * a + 1] <= d_tol;
      const bool pred = conv[4 * a] <= e_pred * conv[4 * a + 2] && conv[4 * a + 1] <= d_pred &&
                        (double)act[a].n * rot * rot <= d_tol;
      if (!last && !pred) next.push_back(act[a]);
    }
    // the finished blocks' vectors into both iterate buffers (the device applies the same test)
    if (next.size() < act.size())
      launch_copy_converged(m_blk, na, m_conv, e_tol, d_tol, e_pred, d_pred, dX.as<double>(), dX2.as<double>(), st);
    act.swap(next);
  }
  oa_unconverged = (int)act.size();
  launch_final_sort(dblk, nb, d_tiles_blk, (int)n_tiles_blk, dX.as<double>(), dlam_tmp.as<double>(),
                    dU.as<double>(), dlam_raw.as<double>(), d_sidx.as<int>(), final_sort_lds(nmax_m), st);
  const int* hs = stg.get(d_stats.as<int>(), 4, st);
  large_blocks_join(*this);
  wide_join();
  sync();
  maxcluster = hs[0];
  unresolved = hs[1];
  large_blocks_check(*this);
  if (need_full_u || fail_on_unconverged) {
    wide_vectors(true);
    sync();
  }
  // A block the refinement left short of its stopping point is solved again in FP64 by the direct
  // solver (the next Hamiltonian is built from these vectors, so they must meet the same bar as every
  // other block's); nrgk_chain_refine_stats reports how many were.  The stand-alone nrgk_eigh_batched
  // fails with NRGK_ERR_NOCONV instead (solve_blocks).
  if (oa_unconverged && !fail_on_unconverged) {
    std::vector<int> secs;
    for (auto& a : act)
      for (int s = 0; s < S; s++)
        if (cur_off[s] == a.off && cur_n[s] == a.n) {
          secs.push_back(s);
          break;
        }
    solve_dense_fp64(*this, secs);
  }
}

// The eigensolver of the current blocks in the chain's mode.  Whatever way it leaves (a status
// thrown from any stage), the side streams that may still write the chain's buffers have drained.
void Chain::eigensolve() {
  try {
    if (mode == NRGK_MODE_PLAIN64) eigensolve_plain();
    else eigensolve_mixed();
  } catch (...) {
    wide_thread_join();
    wide_err = nullptr;
    for (int k = 0; k < N_SIDE; k++)
      if (st_big[k]) (void)hipStreamSynchronize(st_big[k]);
    (void)hipStreamSynchronize(st);
    stg.reset();
    throw;
  }
}

void Chain::add_bath_site(const double* t, int nt, double rescale) {
  NRGK_REQUIRE(phase != ASSEMBLED, NRGK_ERR_STATE,
               "nrgk: add_bath_site called twice without update_internal_state");
  NRGK_REQUIRE(nt == (int)imp.nops, NRGK_ERR_INVALID,
               std::string("t_hopping ") + std::to_string(nt) + " =! pre_fdag_oparator " + std::to_string(imp.nops));
  for (int p = 0; p < nt; p++)
    NRGK_REQUIRE(std::isfinite(t[p]), NRGK_ERR_INVALID, "nrgk: add_bath_site: non-finite hopping");
  NRGK_REQUIRE(std::isfinite(rescale), NRGK_ERR_INVALID, "nrgk: add_bath_site: non-finite rescale");
  step_l0 = thread_counts().launches;
  step_s0 = thread_counts().syncs;
  if (cnt == 0 && phase == FRESH) import_impurity();
  // all_eigenvalue keeps the previous truncation's list until the next update_internal_state
  // (nrgcore.hpp:105-109, 269-327): when it has not been read yet, the raw eigenvalues it is sorted
  // from are kept aside before the eigensolve overwrites them
  if (phase == TRUNCATED && all_eig.empty() && N_tot > 0) {
    dlam_prev.ensure((size_t)N_tot * sizeof(double));
    NRGK_HIP(hipMemcpyAsync(dlam_prev.p, dlam_raw.p, (size_t)N_tot * sizeof(double), hipMemcpyDeviceToDevice, st));
    N_prev = N_tot;
  } else {
    N_prev = 0;
  }
  std::vector<double> tv(t, t + nt);
  NRGK_HIP(hipEventRecord(ev[0], st));
  create_next_basis();
  assemble(tv, rescale);
  NRGK_HIP(hipEventRecord(ev[1], st));
  if (keep_h) {
    h_keep.assign(cur_qn.size(), {});
    for (size_t s = 0; s < cur_qn.size(); s++) {
      h_keep[s].resize((size_t)cur_n[s] * cur_n[s]);
      d2h(h_keep[s].data(), dH.as<double>() + cur_off[s], h_keep[s].size(), st);
    }
    sync();
  }
  eigensolve();
  NRGK_HIP(hipEventRecord(ev[2], st));
  sync();
  NRGK_HIP(hipEventElapsedTime(&ms[0], ev[0], ev[1]));
  NRGK_HIP(hipEventElapsedTime(&ms[1], ev[1], ev[2]));
  phase = ASSEMBLED;
}

// nrgcore::update_internal_state (nrgcore.hpp:105-109)
void Chain::update_internal_state() {
  NRGK_REQUIRE(phase == ASSEMBLED, NRGK_ERR_STATE, "nrgk: update_internal_state without add_bath_site");
  cnt++;
  const int S = (int)cur_qn.size();
  Staging::Region rg = stg.open(Staging::span<int>(cur_n.size()) + Staging::span<int>(cur_eoff.size()));
  const int* d_bn = stg.place(rg, cur_n);
  const int* d_beo = stg.place(rg, cur_eoff);
  stg.commit(rg, st);
  d_trunc.ensure(sizeof(TruncOut) + 256);
  TruncOut* dout = (TruncOut*)d_trunc.p;
  d_kept.ensure(std::max(1, S) * sizeof(int));
  const int lnext = lcur ^ 1;
  dlam_shift[lnext].ensure(std::max<long long>(1, N_tot) * sizeof(double));
  NRGK_HIP(hipEventRecord(ev[3], st));
  launch_truncate(d_bn, d_beo, S, dlam_raw.as<double>(), N_tot, (long long)K, errbar, tie_window, d_kept.as<int>(), dout,
                  dlam_shift[lnext].as<double>(), st);
  const int* h_kept = stg.get(d_kept.as<int>(), S, st);
  const TruncOut* h_out = stg.get(dout, 1, st);
  sync();
  kept.assign(h_kept, h_kept + S);
  const TruncOut out = *h_out;
  last_cut = out.e_cut;
  near_ties = out.near_ties;
  NRGK_REQUIRE(out.nonfinite == 0, NRGK_ERR_NOCONV,
               "nrgk: " + std::to_string(out.nonfinite) + " non-finite eigenvalue(s) in iteration " + std::to_string(cnt));
  // the wide blocks' kept eigenvectors, before the f-dagger rotation reads them
  if (wide_state == WIDE_EIGENVALUES) wide_vectors(false);
  NRGK_HIP(hipEventRecord(ev[4], st));
  all_eig.clear();  // sorted lazily from dlam_raw
  no_kept = (size_t)out.n_kept;
  if (no_kept <= K) cnt_min = cnt;
  else relgs.push_back(out.e_min);
  rotate_fdag();
  NRGK_HIP(hipEventRecord(ev[5], st));
  sync();
  NRGK_HIP(hipEventElapsedTime(&ms[2], ev[3], ev[4]));
  NRGK_HIP(hipEventElapsedTime(&ms[3], ev[4], ev[5]));
  pre_qn = cur_qn;
  pre_kept = kept;
  pre_eoff = cur_eoff;
  lcur = lnext;
  d_eprev = dlam_shift[lcur].as<double>();
  h_last_shift.clear();
  phase = TRUNCATED;
  last_launches = thread_counts().launches - step_l0;
  last_syncs = thread_counts().syncs - step_s0;
}

// f-dagger of the new site in the kept basis (nrgcore::set_current_fdag_operator, nrgcore.hpp:401-477)
void Chain::rotate_fdag() {
  const int S = (int)cur_qn.size();
  const int SB = bath.S, Sp = (int)pre_qn.size();
  // (ii, bb) -> (sector, sub index)
  std::vector<std::pair<int, int>> where((size_t)Sp * SB, {-1, -1});
  for (int s = 0; s < S; s++)
    for (size_t k = 0; k < coupled[s].size(); k++)
      where[(size_t)coupled[s][k].first * SB + coupled[s][k].second] = {s, (int)k};
  std::vector<RotTri> tris;
  std::vector<RotGroup> groups;
  std::vector<RotSegment> segs;
  std::vector<GemmProb> gp;
  std::vector<OpMap> fnew(imp.nops);
  long long yoff = 0, foff = 0;
  // coupled sub-block pairs of one operator: sub-block k of sector i, sub-block sj of sector j, bath
  // block F_bath[p](bb, bbp)
  struct Trip {
    int i, j, k, sj, bbp;
    BlockRef fb;
  };
  std::vector<Trip> tl;
  for (int p = 0; p < imp.nops; p++) {
    // triples (p, i, j) in the reference's map order, (i, j) lexicographic; the sub-block pairs of a
    // triple keep their discovery order (a stable sort on (i, j))
    tl.clear();
    for (int i = 0; i < S; i++)
      for (size_t k = 0; k < coupled[i].size(); k++) {
        auto [ii, bb] = coupled[i][k];
        for (auto& [key, ref] : bathF[p]) {
          if (key.first != bb) continue;
          const auto w = where[(size_t)ii * SB + key.second];
          if (w.first < 0) continue;
          tl.push_back(Trip{i, w.first, (int)k, w.second, key.second, ref});
        }
      }
    std::stable_sort(tl.begin(), tl.end(), [](const Trip& x, const Trip& y) { return x.i != y.i ? x.i < y.i : x.j < y.j; });
    fnew[p].reserve(tl.size());
    for (size_t q0 = 0; q0 < tl.size();) {
      size_t q1 = q0 + 1;
      while (q1 < tl.size() && tl[q1].i == tl[q0].i && tl[q1].j == tl[q0].j) q1++;
      const int i = tl[q0].i, j = tl[q0].j;
      const int ki = kept[i], kj = kept[j], ni = cur_n[i], nj = cur_n[j];
      BlockRef ref{foff, ki, kj};
      fnew[p][{i, j}] = ref;
      foff += (long long)ki * kj;
      if (ki == 0 || kj == 0 || ni == 0) {
        q0 = q1;
        continue;
      }
      const int tri = (int)tris.size();
      tris.push_back(RotTri{cur_off[j], nj, kj, yoff});
      // one group per sub-block of sector i (rows without a coupling stay zero)
      for (size_t k = 0; k < coupled[i].size(); k++) {
        auto [ii, bb] = coupled[i][k];
        const int kii = pre_kept[ii], dbb = bath.dim(bb);
        if (kii * dbb == 0) continue;
        RotGroup g{tri, sub_kidx[i][k], kii, dbb, (int)segs.size(), 0};
        for (size_t q = q0; q < q1; q++) {
          if (tl[q].k != (int)k) continue;
          segs.push_back(RotSegment{sub_kidx[j][tl[q].sj], bath.dim(tl[q].bbp), (int)tl[q].fb.off, tl[q].fb.cols});
        }
        g.seg_end = (int)segs.size();
        groups.push_back(g);
      }
      gp.push_back(GemmProb{dU.as<double>() + cur_off[i], nullptr, nullptr, nullptr, ki, kj, ni, ni, kj, kj, 0, 1, 0,
                            1.0, 0.0, 0, 0});
      gp.back().B = (const double*)(intptr_t)yoff;  // patched below once Y is allocated
      gp.back().C = (double*)(intptr_t)ref.off;
      yoff += (long long)ni * kj;
      q0 = q1;
    }
  }
  const int nf = fcur ^ 1;
  fbuf[nf].ensure(std::max<long long>(1, foff) * sizeof(double));
  dY.ensure(std::max<long long>(1, yoff) * sizeof(double));
  for (auto& g : gp) {
    g.B = dY.as<double>() + (intptr_t)g.B;
    g.C = fbuf[nf].as<double>() + (intptr_t)g.C;
  }
  Staging::Region rg = stg.open(Staging::span<RotGroup>(groups.size()) + Staging::span<RotSegment>(segs.size()) +
                                Staging::span<RotTri>(tris.size()));
  const RotGroup* d_groups = stg.place(rg, groups);
  const RotSegment* d_segs = stg.place(rg, segs);
  const RotTri* d_tris = stg.place(rg, tris);
  stg.commit(rg, st);
  launch_fdag_gather(d_groups, (int)groups.size(), d_segs, d_tris, dU.as<double>(), d_bathF.as<double>(),
                     dY.as<double>(), st);
  run_gemms(gp);
  fprev = std::move(fnew);
  fcur = nf;
}

void Chain::ensure_all_eig() {
  if (!all_eig.empty() || phase == FRESH) return;
  // after add_bath_site: the previous iteration's eigenvalues, kept aside there
  const bool prev = phase == ASSEMBLED;
  const long long n = prev ? N_prev : N_tot;
  if (n == 0) return;
  std::vector<double> v(n);
  d2h(v.data(), (prev ? dlam_prev : dlam_raw).as<double>(), n, st);
  sync();
  std::sort(v.begin(), v.end());
  all_eig = std::move(v);
}

// NrgData::loadCurrentData-style state overwrite (nrgData.hpp:266-293): the previous sectors, their
// kept energies and the kept x kept f-dagger blocks the next assembly reads.
void Chain::set_state(const nrgk_site_desc* d, int cnt_) {
  SiteData s = SiteData::from_desc(d);
  NRGK_REQUIRE(s.nqn == imp.nqn && s.nops == imp.nops, NRGK_ERR_INVALID, "nrgk: state does not match the model");
  pre_qn = s.qn;
  pre_kept.assign(s.S, 0);
  pre_eoff.assign(s.S, 0);
  std::vector<double> flat;
  for (int i = 0; i < s.S; i++) {
    pre_kept[i] = s.dim(i);
    pre_eoff[i] = (int)flat.size();
    flat.insert(flat.end(), s.E[i].begin(), s.E[i].end());
  }
  d_eprev = upload(eprev_own, flat, st);
  std::vector<double> f;
  fprev.assign(s.nops, OpMap());
  for (int p = 0; p < s.nops; p++)
    for (auto& [key, blk] : s.ops[p]) {
      fprev[p][key] = BlockRef{(long long)f.size(), s.dim(key.first), s.dim(key.second)};
      f.insert(f.end(), blk.begin(), blk.end());
    }
  upload(fbuf[fcur], f, st);
  sync();
  cnt = cnt_;
  phase = TRUNCATED;
  h_last_shift = flat;
}

}  // namespace nrgk
