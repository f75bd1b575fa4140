// Observables along the chain, the T = 0 FDM backward pass and the thermodynamics.
//
//   Obs::update   restates update_system_operator (nrgcore/include/nrgcore/sysOperator.hpp:35-133)
//   Fdm::record   keeps in HBM what NrgData::saveCurrentData / saveqOperator write to an HDF5 temp
//                 file (nrgcore/include/nrgcore/nrgData.hpp:134-205)
//   Fdm::backward restates fdmSpectrum::calcSpectrum over the saved shells
//                 (nrgcore/include/dynamics/fdmSpectrum.hpp:55-63, 71-114, 122-199, 207-277, 289-298)
//   Fdm::thermo   restates fdmThermodynamics (nrgcore/include/static/fdmThermodynamics.hpp:42-236)
#include "observables.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "h5io.hpp"
#include "kernels.hpp"

namespace nrgk {

Obs::Obs(Chain* c, int n_ops, const int* nblocks, const int* bi, const int* bj, const double* data) : ch(c), nops(n_ops) {
  for (DevBuf* b : {&buf[0], &buf[1], &dseg, &dC, &dY}) b->ctx = ch->ctx;
  NRGK_REQUIRE(ch->phase == Chain::FRESH, NRGK_ERR_STATE, "nrgk: observables are created before the first site");
  ch->need_full_u = true;  // the rotation reads every column of U
  blocks.assign(nops, OpMap());
  std::vector<double> f;
  size_t bo = 0, dof = 0;
  for (int p = 0; p < nops; p++)
    for (int b = 0; b < nblocks[p]; b++, bo++) {
      const int i = bi[bo], j = bj[bo];
      NRGK_REQUIRE(i >= 0 && i < ch->imp.S && j >= 0 && j < ch->imp.S, NRGK_ERR_INVALID, "nrgk: observable block out of range");
      const int r = ch->imp.dim(i), cc = ch->imp.dim(j);
      NRGK_REQUIRE(!blocks[p].count({i, j}), NRGK_ERR_INVALID, "Found operator already");
      blocks[p][{i, j}] = BlockRef{(long long)f.size(), r, cc};
      f.insert(f.end(), data + dof, data + dof + (size_t)r * cc);
      dof += (size_t)r * cc;
    }
  buf[cur].ensure(std::max<size_t>(1, f.size()) * sizeof(double));
  h2d(buf[cur].as<double>(), f.data(), f.size(), ch->st);
  NRGK_HIP(stream_sync(ch->st));
}

void Obs::update() {
  Chain& c = *ch;
  NRGK_REQUIRE(c.phase == Chain::ASSEMBLED, NRGK_ERR_STATE, "nrgk: update_system_operator needs add_bath_site first");
  const int S = (int)c.cur_qn.size();
  const int nb = c.bath.S;
  std::vector<OpMap> nbk(nops);
  std::vector<ObsSeg> segs;
  std::vector<GemmProb> g1, g2;
  long long coff = 0, ooff = 0;
  struct Pair {
    int p, i, j;
    long long c_off, o_off;
  };
  std::vector<Pair> pairs;
  // index: previous sector ii -> list of (sector, sub index) that contain it
  std::vector<std::vector<std::pair<int, int>>> by_ii(c.pre_qn.size());
  for (int s = 0; s < S; s++)
    for (size_t k = 0; k < c.coupled[s].size(); k++) by_ii[c.coupled[s][k].first].push_back({s, (int)k});
  struct Item {
    int i, j, k, kp;
    long long o_off;
    int o_cols;
  };
  std::vector<Item> items;
  for (int p = 0; p < nops; p++) {
    // (i, j) pairs in sector order; a pair exists when some (ii,bb) in i and (ii',bb) in j have
    // O(ii,ii').  The (k, kp) of a pair keep their discovery order (a stable sort on (i, j)).
    items.clear();
    for (auto& [key, ref] : blocks[p]) {
      const int ii = key.first, iip = key.second;
      if (ii >= (int)by_ii.size() || iip >= (int)by_ii.size()) continue;
      for (auto [i, k] : by_ii[ii])
        for (auto [j, kp] : by_ii[iip])
          if (c.coupled[i][k].second == c.coupled[j][kp].second)
            items.push_back(Item{i, j, k, kp, ref.off, ref.cols});
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const Item& x, const Item& y) { return x.i != y.i ? x.i < y.i : x.j < y.j; });
    (void)nb;
    for (size_t q0 = 0; q0 < items.size();) {
      size_t q1 = q0 + 1;
      while (q1 < items.size() && items[q1].i == items[q0].i && items[q1].j == items[q0].j) q1++;
      const int i = items[q0].i, j = items[q0].j;
      const int ni = c.cur_n[i], nj = c.cur_n[j];
      Pair pr{p, i, j, coff, ooff};
      for (size_t q = q0; q < q1; q++) {
        const int k = items[q].k, kp = items[q].kp;
        const int ii = c.coupled[i][k].first, bb = c.coupled[i][k].second;
        const int iip = c.coupled[j][kp].first;
        const int kii = c.pre_kept[ii], kiip = c.pre_kept[iip], dbb = c.bath.dim(bb);
        if ((long long)kii * kiip * dbb == 0) continue;
        segs.push_back(ObsSeg{coff, nj, c.sub_kidx[i][k], c.sub_kidx[j][kp], kii, kiip, dbb, items[q].o_off,
                              items[q].o_cols});
      }
      nbk[p][{i, j}] = BlockRef{ooff, ni, nj};
      coff += (long long)ni * nj;
      ooff += (long long)ni * nj;
      pairs.push_back(pr);
      q0 = q1;
    }
  }
  const int nbuf = cur ^ 1;
  buf[nbuf].ensure(std::max<long long>(1, ooff) * sizeof(double));
  dC.ensure(std::max<long long>(1, coff) * sizeof(double));
  dY.ensure(std::max<long long>(1, coff) * sizeof(double));
  NRGK_HIP(hipMemsetAsync(dC.p, 0, std::max<long long>(1, coff) * sizeof(double), c.st));
  const ObsSeg* d_segs = c.stg.put(segs, c.st);
  launch_obs_scatter(d_segs, (int)segs.size(), buf[cur].as<double>(), dC.as<double>(), c.st);
  for (auto& pr : pairs) {
    const int ni = c.cur_n[pr.i], nj = c.cur_n[pr.j];
    double* Cm = dC.as<double>() + pr.c_off;
    double* Y = dY.as<double>() + pr.c_off;
    double* O = buf[nbuf].as<double>() + pr.o_off;
    const double* Ui = c.dU.as<double>() + c.cur_off[pr.i];
    const double* Uj = c.dU.as<double>() + c.cur_off[pr.j];
    // Y = c U_j ; O = U_i^T Y   (sysOperator.hpp:121-122)
    g1.push_back(GemmProb{Cm, Uj, nullptr, Y, ni, nj, nj, nj, nj, nj, 0, 0, 0, 1.0, 0.0, 0, 0});
    g2.push_back(GemmProb{Ui, Y, nullptr, O, ni, nj, ni, ni, nj, nj, 0, 1, 0, 1.0, 0.0, 0, 0});
  }
  c.run_gemms(g1);
  c.run_gemms(g2);
  c.sync();
  blocks = std::move(nbk);
  cur = nbuf;
}

// ------------------------------------------------------------------ FDM
Fdm::Fdm(Chain* c, Obs* o, int nb, double emin, double emax) : ch(c), obs(o), nbins(nb), e_min(emin), e_max(emax) {
  NRGK_REQUIRE(nb > 0 && emin > 0 && emax > emin, NRGK_ERR_INVALID, "nrgk: bad FDM binning");
  delE = (nbins - 1.0) / std::log(e_max / e_min);  // fdmSpectrum.hpp:351-354
  for (DevBuf* b : {&dpos, &dneg, &dsum, &dbf, &dred[0], &dred[1], &dRA, &dAR, &dT1, &drho, &ditems, &dpt, &dE, &dth, &dQ})
    b->ctx = ch->ctx;
}

void* Fdm::arena_alloc(size_t bytes) {
  bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
  if (chunks.empty() || chunk_used + bytes > chunks.back()->cap) {
    auto b = std::make_unique<DevBuf>(ch->ctx);
    b->ensure(std::max<size_t>(bytes, size_t(512) << 20));
    chunks.push_back(std::move(b));
    chunk_used = 0;
  }
  void* p = static_cast<char*>(chunks.back()->p) + chunk_used;
  chunk_used += bytes;
  return p;
}

void Fdm::record() {
  Chain& c = *ch;
  NRGK_REQUIRE(c.phase == Chain::ASSEMBLED, NRGK_ERR_STATE, "nrgk: record after add_bath_site, before update_internal_state");
  auto sh = std::make_unique<Shell>();
  sh->cnt = c.cnt;
  sh->n = c.cur_n;
  sh->eoff = c.cur_eoff;
  sh->uoff = c.cur_off;
  sh->coupled = c.coupled;
  sh->sub_kidx = c.sub_kidx;
  sh->pre_kept = c.pre_kept;
  sh->qn = c.cur_qn;
  sh->pre_qn = c.pre_qn;
  sh->n_bath = c.bath.S;
  const double* h_e = c.stg.get(c.dlam_raw.as<double>(), c.N_tot, c.st);
  if (obs) {
    sh->dU.p = arena_alloc(std::max<long long>(1, c.tot_elems) * sizeof(double));
    NRGK_HIP(hipMemcpyAsync(sh->dU.p, c.dU.p, c.tot_elems * sizeof(double), hipMemcpyDeviceToDevice, c.st));
    sh->dE.p = arena_alloc(std::max<long long>(1, c.N_tot) * sizeof(double));
    NRGK_HIP(hipMemcpyAsync(sh->dE.p, c.dlam_raw.p, c.N_tot * sizeof(double), hipMemcpyDeviceToDevice, c.st));
    long long tot = 0;
    for (auto& m : obs->blocks)
      for (auto& [k, r] : m) tot = std::max(tot, r.off + (long long)r.rows * r.cols);
    sh->dO.p = arena_alloc(std::max<long long>(1, tot) * sizeof(double));
    NRGK_HIP(hipMemcpyAsync(sh->dO.p, obs->buf[obs->cur].p, tot * sizeof(double), hipMemcpyDeviceToDevice, c.st));
    sh->ops = obs->blocks;
  }
  // the system operators' diagonals in this shell's eigenbasis (fdmThermodynamics.hpp:63-87)
  const double* h_q = nullptr;
  if (sysop) {
    const int S = (int)c.cur_n.size();
    std::vector<DiagItem> items;
    for (int q = 0; q < sysop->nops; q++)
      for (int i = 0; i < S; i++) {
        auto f = sysop->blocks[q].find({i, i});
        NRGK_REQUIRE(f == sysop->blocks[q].end() || (f->second.rows == c.cur_n[i] && f->second.cols == c.cur_n[i]),
                     NRGK_ERR_STATE, "nrgk: system operator not updated for this iteration");
        items.push_back({f == sysop->blocks[q].end() ? -1LL : f->second.off, (long long)q * c.N_tot + c.cur_eoff[i],
                         c.cur_n[i]});
      }
    const size_t nq = (size_t)sysop->nops * c.N_tot;
    dQ.ensure(std::max<size_t>(1, nq) * sizeof(double));
    launch_diag_gather(c.stg.put(items, c.st), (int)items.size(), sysop->buf[sysop->cur].as<double>(),
                       dQ.as<double>(), c.st);
    h_q = c.stg.get(dQ.as<double>(), nq, c.st);
    sh->q_ops = sysop->nops;
  }
  c.sync();
  sh->E.assign(h_e, h_e + c.N_tot);
  if (h_q) sh->Q.assign(h_q, h_q + (size_t)sh->q_ops * c.N_tot);
  shells.push_back(std::move(sh));
}

// The reference's temporary file of the forward pass (NrgData, nrgData.hpp:134-205) for every recorded
// shell: /NrgItr{cnt}/ with current_hamiltonQ{s} (the n_s x n_s eigenvector block, row-major, columns
// the eigenvectors), current_sysmQ{s}, pre_sysmQ{s}, eigenvaluesQ{s}, coupled_nQ_index{s}
// (i * bath sectors + j) and eigenvaluesQ_kept_indices{s}; Eigenvalues{cnt} at the root; and the
// observable as saveqOperator writes it, in the group {op_group}{cnt - 1} (the SIAM drivers' index,
// examples/fdmSpectrumSiam/main.cpp:89-90): attribute operatorSize, at{p}/qmat{k}, idVector{k},
// colVector, rowVector.  Eigenvalues{cnt} holds the shell's own eigenvalues, ascending (the reference
// writes the chain's all_eigenvalue of that moment, the previous iteration's list).
void Fdm::save(const std::string& path, const std::string& op_group) const {
  Chain& c = *ch;
  h5::File f(path, true);
  std::vector<double> hU, hO;
  for (const auto& shp : shells) {
    const Shell& sh = *shp;
    const std::string g = "/NrgItr" + std::to_string(sh.cnt) + "/";
    f.group(g);
    const int S = (int)sh.n.size();
    if (sh.dU.p) {
      long long tot = 0;
      for (int i = 0; i < S; i++) tot = std::max(tot, sh.uoff[i] + (long long)sh.n[i] * sh.n[i]);
      hU.resize(std::max<long long>(1, tot));
      d2h(hU.data(), sh.dU.as<double>(), tot, c.st);
    }
    if (sh.dO.p) {
      long long tot = 0;
      for (auto& m : sh.ops)
        for (auto& [k, r] : m) tot = std::max(tot, r.off + (long long)r.rows * r.cols);
      hO.resize(std::max<long long>(1, tot));
      d2h(hO.data(), sh.dO.as<double>(), tot, c.st);
    }
    NRGK_HIP(stream_sync(c.st));
    for (int i = 0; i < S; i++) {
      const std::string si = std::to_string(i);
      if (sh.dU.p)
        f.write(g + "current_hamiltonQ" + si, h5::Type::F64, hU.data() + sh.uoff[i], (size_t)sh.n[i] * sh.n[i]);
      f.write(g + "current_sysmQ" + si, sh.qn[i]);
      f.write(g + "eigenvaluesQ" + si, std::vector<double>(sh.E.begin() + sh.eoff[i], sh.E.begin() + sh.eoff[i] + sh.n[i]));
      std::vector<unsigned long> cq;
      for (auto [ii, bb] : sh.coupled[i]) cq.push_back((unsigned long)ii * sh.n_bath + bb);
      f.write(g + "coupled_nQ_index" + si, cq);
    }
    for (size_t i = 0; i < sh.pre_qn.size(); i++) {
      const std::string si = std::to_string(i);
      f.write(g + "pre_sysmQ" + si, sh.pre_qn[i]);
      std::vector<unsigned long> kept(sh.pre_kept[i]);
      for (int k = 0; k < sh.pre_kept[i]; k++) kept[k] = k;
      f.write(g + "eigenvaluesQ_kept_indices" + si, kept);
    }
    std::vector<double> all = sh.E;
    std::sort(all.begin(), all.end());
    f.write("Eigenvalues" + std::to_string(sh.cnt), all);
    if (sh.dO.p && !op_group.empty()) {
      const std::string og = op_group + std::to_string(sh.cnt - 1);
      f.attribute(og, "operatorSize", (unsigned long)sh.ops.size());
      for (size_t p = 0; p < sh.ops.size(); p++) {
        const std::string lg = og + "/at" + std::to_string(p) + "/";
        f.group(lg);
        std::vector<unsigned long> cols, rows;
        size_t k = 0;
        for (auto& [key, r] : sh.ops[p]) {
          f.write(lg + "qmat" + std::to_string(k), h5::Type::F64, hO.data() + r.off, (size_t)r.rows * r.cols);
          f.write(lg + "idVector" + std::to_string(k), std::vector<unsigned long>{(unsigned long)key.first, (unsigned long)key.second});
          cols.push_back(r.cols);
          rows.push_back(r.rows);
          k++;
        }
        f.write(lg + "colVector", cols);
        f.write(lg + "rowVector", rows);
      }
    }
  }
}

// NrgData::loadCurrentData + loadqOperator (nrgData.hpp:214-293) for every shell of a file that
// Fdm::save wrote: the shells are rebuilt on the device (U blocks, eigenvalues and, when op_group is
// given and the recorder has an observable, the operator blocks), so that a backward pass or the
// thermodynamics can run in another process than the forward pass.  The chain must have the bath the
// file was written with (its sector dimensions place the sub-blocks).  Shells already recorded are
// replaced.
void Fdm::load(const std::string& path, const std::string& op_group) {
  Chain& c = *ch;
  h5::File f(path, false);
  int first = -1;
  for (int k = 0; k < 100000 && first < 0; k++)
    if (f.exists("/NrgItr" + std::to_string(k))) first = k;
  NRGK_REQUIRE(first >= 0, NRGK_ERR_INVALID, "nrgk_fdm_load: no /NrgItr group in " + path);
  shells.clear();
  chunks.clear();
  chunk_used = 0;
  const int nb = c.bath.S;
  for (int cnt = first; f.exists("/NrgItr" + std::to_string(cnt)); cnt++) {
    const std::string g = "/NrgItr" + std::to_string(cnt) + "/";
    auto sh = std::make_unique<Shell>();
    sh->cnt = cnt;
    sh->n_bath = nb;
    for (int i = 0; f.exists(g + "pre_sysmQ" + std::to_string(i)); i++) {
      const std::string si = std::to_string(i);
      std::vector<int> q(f.size(g + "pre_sysmQ" + si));
      f.read(g + "pre_sysmQ" + si, h5::Type::I32, q.data(), q.size());
      sh->pre_qn.push_back(q);
      sh->pre_kept.push_back((int)f.size(g + "eigenvaluesQ_kept_indices" + si));
    }
    long long uo = 0;
    for (int i = 0; f.exists(g + "current_sysmQ" + std::to_string(i)); i++) {
      const std::string si = std::to_string(i);
      std::vector<int> q(f.size(g + "current_sysmQ" + si));
      f.read(g + "current_sysmQ" + si, h5::Type::I32, q.data(), q.size());
      sh->qn.push_back(q);
      const int n = (int)f.size(g + "eigenvaluesQ" + si);
      sh->eoff.push_back((int)sh->E.size());
      sh->n.push_back(n);
      sh->uoff.push_back(uo);
      uo += (long long)n * n;
      std::vector<double> e(n);
      f.read(g + "eigenvaluesQ" + si, h5::Type::F64, e.data(), n);
      sh->E.insert(sh->E.end(), e.begin(), e.end());
      std::vector<unsigned long> cq(f.size(g + "coupled_nQ_index" + si));
      f.read(g + "coupled_nQ_index" + si, h5::Type::U64, cq.data(), cq.size());
      std::vector<std::pair<int, int>> cp;
      std::vector<int> kidx;
      int k0 = 0;
      for (unsigned long v : cq) {
        const int ii = (int)(v / nb), bb = (int)(v % nb);
        NRGK_REQUIRE(ii < (int)sh->pre_kept.size(), NRGK_ERR_INVALID, "nrgk_fdm_load: coupled index out of range");
        cp.push_back({ii, bb});
        kidx.push_back(k0);
        k0 += sh->pre_kept[ii] * c.bath.dim(bb);
      }
      NRGK_REQUIRE(k0 == n, NRGK_ERR_INVALID, "nrgk_fdm_load: the chain's bath does not match the file");
      sh->coupled.push_back(cp);
      sh->sub_kidx.push_back(kidx);
    }
    const int S = (int)sh->n.size();
    if (obs && f.exists(g + "current_hamiltonQ0")) {
      std::vector<double> hU(std::max<long long>(1, uo));
      for (int i = 0; i < S; i++)
        f.read(g + "current_hamiltonQ" + std::to_string(i), h5::Type::F64, hU.data() + sh->uoff[i],
               (size_t)sh->n[i] * sh->n[i]);
      sh->dU.p = arena_alloc(hU.size() * sizeof(double));
      h2d(sh->dU.as<double>(), hU.data(), hU.size(), c.st);
      sh->dE.p = arena_alloc(std::max<size_t>(1, sh->E.size()) * sizeof(double));
      h2d(sh->dE.as<double>(), sh->E.data(), sh->E.size(), c.st);
      NRGK_REQUIRE(!op_group.empty(), NRGK_ERR_INVALID, "nrgk_fdm_load: the recorder's observable needs op_group");
      const std::string og = op_group + std::to_string(cnt - 1);
      sh->ops.assign(obs->nops, OpMap());
      std::vector<double> hO;
      for (int p = 0; p < obs->nops; p++) {
        const std::string lg = og + "/at" + std::to_string(p) + "/";
        NRGK_REQUIRE(f.exists(lg + "rowVector"), NRGK_ERR_INVALID, "nrgk_fdm_load: no operator " + lg + " in " + path);
        std::vector<unsigned long> rows(f.size(lg + "rowVector")), cols(f.size(lg + "colVector"));
        f.read(lg + "rowVector", h5::Type::U64, rows.data(), rows.size());
        f.read(lg + "colVector", h5::Type::U64, cols.data(), cols.size());
        for (size_t k = 0; k < rows.size(); k++) {
          unsigned long id[2];
          f.read(lg + "idVector" + std::to_string(k), h5::Type::U64, id, 2);
          const size_t sz = (size_t)rows[k] * cols[k];
          const long long off = (long long)hO.size();
          hO.resize(hO.size() + sz);
          f.read(lg + "qmat" + std::to_string(k), h5::Type::F64, hO.data() + off, sz);
          sh->ops[p][{(int)id[0], (int)id[1]}] = BlockRef{off, (int)rows[k], (int)cols[k]};
        }
      }
      sh->dO.p = arena_alloc(std::max<size_t>(1, hO.size()) * sizeof(double));
      h2d(sh->dO.as<double>(), hO.data(), hO.size(), c.st);
      NRGK_HIP(stream_sync(c.st));
    }
    shells.push_back(std::move(sh));
  }
  // the count the forward pass ended at (after the last shell's update_internal_state): the shell
  // prefactors of the thermodynamics count from it
  loaded_nrg_max = shells.back()->cnt + 1;
}

// the recorded shell with iteration counter cnt
Shell* Fdm::shell_at(int cnt) const {
  for (auto& s : shells)
    if (s->cnt == cnt) return s.get();
  throw Error(NRGK_ERR_STATE, "nrgk: no shell was recorded at iteration " + std::to_string(cnt));
}

void Fdm::backward(double lambda, int scale_offset, int first_cnt, double* pos, double* neg, double* traces,
                   double* sums, int cap, int* n_shells) {
  std::vector<std::pair<Shell*, double>> use;
  for (auto it = shells.rbegin(); it != shells.rend(); ++it)
    if ((*it)->cnt >= first_cnt) use.push_back({it->get(), std::pow(lambda, -((double)(*it)->cnt + scale_offset) / 2.0)});
  backward_shells(use, pos, neg, traces, sums, cap, n_shells);
}

// The backward pass over the given shells in the given order (the last shell of the chain first),
// each with its energy scale: the loop of calcSpectrum(scale) calls (fdmSpectrum.hpp:55-63) of the
// reference's drivers (examples/fdmSpectrumSiam/main.cpp:158-180).
void Fdm::backward_shells(const std::vector<std::pair<Shell*, double>>& list, double* pos, double* neg, double* traces,
                          double* sums, int cap, int* n_shells) {
  NRGK_REQUIRE(obs, NRGK_ERR_STATE, "nrgk: the FDM spectrum needs an observable");
  Chain& c = *ch;
  hipStream_t st = c.st;
  const int nops = obs->nops;
  dpos.ensure((size_t)nops * nbins * sizeof(double));
  dneg.ensure((size_t)nops * nbins * sizeof(double));
  // [0]: this shell's sum of |w|; [1 + p]: operator p's signed weights over all shells
  dsum.ensure(sizeof(double) * (1 + nops));
  NRGK_HIP(hipMemsetAsync(dsum.p, 0, sizeof(double) * (1 + nops), st));
  NRGK_HIP(hipMemsetAsync(dpos.p, 0, (size_t)nops * nbins * sizeof(double), st));
  NRGK_HIP(hipMemsetAsync(dneg.p, 0, (size_t)nops * nbins * sizeof(double), st));
  std::vector<Shell*> use;
  for (auto& u : list) use.push_back(u.first);
  int ns = 0;
  std::vector<int> kc;               // kept counts of the shell being processed (currentKeptIndex)
  std::vector<long long> red_off;    // reducedRho offsets (indexed by this shell's sectors)
  int rcur = 0;
  for (size_t u = 0; u < use.size(); u++) {
    Shell& sh = *use[u];
    const int S = (int)sh.n.size();
    const bool last = (u == 0);
    const double scale = list[u].second;
    // currentKeptIndex (fdmSpectrum.hpp:289-298): empty at the last shell, else the kept set saved
    // with the later shell
    if (last) kc.assign(S, 0);
    else {
      kc = use[u - 1]->pre_kept;
      NRGK_REQUIRE((int)kc.size() == S, NRGK_ERR_STATE, "nrgk: recorded shells are not consecutive");
    }
    double trace = 0.0;
    // rho (setRhoZero, fdmSpectrum.hpp:165-199)
    std::vector<double> bf;
    if (last) {
      // setLocalPartitionFunction (fdmSpectrum.hpp:122-155)
      double egs = 0.0;
      for (int i = 0; i < S; i++)
        if (sh.n[i] > 0) egs = std::min(egs, *std::min_element(sh.E.begin() + sh.eoff[i], sh.E.begin() + sh.eoff[i] + sh.n[i]));
      double Z = 0.0;
      for (double e : sh.E)
        if (std::fabs(e - egs) < 1e-5) Z += 1.;
      bf.assign(sh.E.size(), 0.0);
      for (size_t e = 0; e < sh.E.size(); e++)
        if (std::fabs(sh.E[e] - egs) < 1e-5) bf[e] = 1. / Z;
      for (double x : bf) trace += x;
      dbf.ensure(std::max<size_t>(1, bf.size()) * sizeof(double));
      h2d(dbf.as<double>(), bf.data(), bf.size(), st);
    } else {
      long long tot = 0;
      for (int i = 0; i < S; i++) tot += (long long)kc[i] * kc[i];
      std::vector<double> hred(tot);
      d2h(hred.data(), dred[rcur].as<double>(), tot, st);
      NRGK_HIP(stream_sync(st));
      for (int i = 0; i < S; i++)
        for (int a = 0; a < kc[i]; a++) trace += hred[red_off[i] + (long long)a * kc[i] + a];
    }
    // spectral weights (rhoDotOperators, fdmSpectrum.hpp:207-277)
    std::vector<FdmBinItem> items;
    std::vector<GemmProb> g;
    long long ra = 0, ar = 0;
    const double* red = dred[rcur].as<double>();
    for (int p = 0; p < nops; p++)
      for (auto& [ij, ref] : sh.ops[p]) {
        const int i = ij.first, j = ij.second;
        const int ni = sh.n[i], nj = sh.n[j];
        if (ni == 0 || nj == 0) continue;
        FdmBinItem it{};
        it.b_off = ref.off;
        it.ni = ni;
        it.nj = nj;
        it.kci = kc[i];
        it.kcj = kc[j];
        it.ei = sh.eoff[i];
        it.ej = sh.eoff[j];
        it.op = p;
        it.diag_mode = last ? 1 : 0;
        if (!last) {
          it.ra_off = ra;
          it.ar_off = ar;
          ra += (long long)ni * kc[j];
          ar += (long long)kc[i] * nj;
        }
        const int rows_per = std::max(1, 16384 / nj);
        for (int r0 = 0; r0 < ni; r0 += rows_per) {
          FdmBinItem t = it;
          t.row0 = r0;
          t.rows = std::min(rows_per, ni - r0);
          items.push_back(t);
        }
      }
    dRA.ensure(std::max<long long>(1, ra) * sizeof(double));
    dAR.ensure(std::max<long long>(1, ar) * sizeof(double));
    if (!last) {
      // rhoA = B[:, :kc_j] rho_j ; ARho = rho_i B[:kc_i, :]
      long long rao = 0, aro = 0;
      for (int p = 0; p < nops; p++)
        for (auto& [ij, ref] : sh.ops[p]) {
          const int i = ij.first, j = ij.second;
          const int ni = sh.n[i], nj = sh.n[j];
          if (ni == 0 || nj == 0) continue;
          const double* B = sh.dO.as<double>() + ref.off;
          if (kc[j] > 0)
            g.push_back(GemmProb{B, red + red_off[j], nullptr, dRA.as<double>() + rao, ni, kc[j], kc[j], nj, kc[j], kc[j],
                                 0, 0, 0, 1.0, 0.0, 0, 0});
          if (kc[i] > 0)
            g.push_back(GemmProb{red + red_off[i], B, nullptr, dAR.as<double>() + aro, kc[i], nj, kc[i], kc[i], nj, nj, 0,
                                 0, 0, 1.0, 0.0, 0, 0});
          rao += (long long)ni * kc[j];
          aro += (long long)kc[i] * nj;
        }
      c.run_gemms(g);
    }
    NRGK_HIP(hipMemsetAsync(dsum.p, 0, sizeof(double), st));
    ditems.ensure(std::max<size_t>(1, items.size()) * sizeof(FdmBinItem));
    h2d(ditems.as<FdmBinItem>(), items.data(), items.size(), st);
    // algorithmic bytes (the binning's roofline): B, rhoA and ARho read once per (v, v') pair
    double bin_bytes = 0;
    if (c.prof_on)
      for (const auto& it : items) bin_bytes += 24.0 * (double)it.rows * it.nj;
    c.prof_begin("fdm_bin");
    launch_fdm_bin(ditems.p, (int)items.size(), sh.dO.as<double>(), dRA.as<double>(), dAR.as<double>(),
                   sh.dE.as<double>(), last ? dbf.as<double>() : nullptr, scale, e_min, delE, nbins, dpos.as<double>(),
                   dneg.as<double>(), dsum.as<double>(), dsum.as<double>() + 1, st);
    c.prof_end("fdm_bin", 1, 0.0, bin_bytes);
    double ssum = 0;
    d2h(&ssum, dsum.as<double>(), 1, st);
    // reduced density matrix (setReduceDensityMatrix, fdmSpectrum.hpp:71-114):
    // rho_i <- U_i rho_i U_i^T, then the partial trace over the last site
    long long t1 = 0, rf = 0;
    std::vector<long long> t1off(S), rfoff(S);
    for (int i = 0; i < S; i++) {
      const int k = last ? sh.n[i] : kc[i];
      t1off[i] = t1;
      rfoff[i] = rf;
      t1 += (long long)sh.n[i] * k;
      rf += (long long)sh.n[i] * sh.n[i];
    }
    dT1.ensure(std::max<long long>(1, t1) * sizeof(double));
    drho.ensure(std::max<long long>(1, rf) * sizeof(double));
    NRGK_HIP(hipMemsetAsync(drho.p, 0, std::max<long long>(1, rf) * sizeof(double), st));
    std::vector<GemmProb> g1, g2;
    DevBuf ddiag(c.ctx);
    if (last) {
      // rho is diagonal: T1 = U diag(bf) as a GEMM with a dense diagonal matrix
      long long dd = 0;
      for (int i = 0; i < S; i++) dd += (long long)sh.n[i] * sh.n[i];
      std::vector<double> dm(std::max<long long>(1, dd), 0.0);
      for (int i = 0; i < S; i++)
        for (int a = 0; a < sh.n[i]; a++) dm[rfoff[i] + (long long)a * sh.n[i] + a] = bf[sh.eoff[i] + a];
      ddiag.ensure(dm.size() * sizeof(double));
      h2d(ddiag.as<double>(), dm.data(), dm.size(), st);
      for (int i = 0; i < S; i++) {
        const int n = sh.n[i];
        if (n == 0) continue;
        const double* U = sh.dU.as<double>() + sh.uoff[i];
        g1.push_back(GemmProb{U, ddiag.as<double>() + rfoff[i], nullptr, dT1.as<double>() + t1off[i], n, n, n, n, n, n, 0,
                              0, 0, 1.0, 0.0, 0, 0});
        g2.push_back(GemmProb{dT1.as<double>() + t1off[i], U, nullptr, drho.as<double>() + rfoff[i], n, n, n, n, n, n, 0,
                              0, 1, 1.0, 0.0, 0, 0});
      }
    } else {
      for (int i = 0; i < S; i++) {
        const int n = sh.n[i], k = kc[i];
        if (n == 0 || k == 0) continue;
        const double* U = sh.dU.as<double>() + sh.uoff[i];
        g1.push_back(GemmProb{U, red + red_off[i], nullptr, dT1.as<double>() + t1off[i], n, k, k, n, k, k, 0, 0, 0, 1.0,
                              0.0, 0, 0});
        g2.push_back(GemmProb{dT1.as<double>() + t1off[i], U, nullptr, drho.as<double>() + rfoff[i], n, n, k, k, n, n, 0,
                              0, 1, 1.0, 0.0, 0, 0});
      }
    }
    c.run_gemms(g1);
    c.run_gemms(g2);
    // partial trace into reducedRho of the previous sectors
    const int Sp = (int)sh.pre_kept.size();
    std::vector<long long> new_off(Sp);
    long long ro = 0;
    for (int ii = 0; ii < Sp; ii++) {
      new_off[ii] = ro;
      ro += (long long)sh.pre_kept[ii] * sh.pre_kept[ii];
    }
    std::vector<std::vector<PtContrib>> per(Sp);
    for (int i = 0; i < S; i++)
      for (size_t k = 0; k < sh.coupled[i].size(); k++) {
        auto [ii, bb] = sh.coupled[i][k];
        const int kii = sh.pre_kept[ii];
        if (kii == 0 || sh.n[i] == 0) continue;
        per[ii].push_back(PtContrib{rfoff[i], sh.n[i], sh.sub_kidx[i][k], c.bath.dim(bb)});
      }
    std::vector<PtItem> pit;
    std::vector<PtContrib> pcs;
    for (int ii = 0; ii < Sp; ii++) {
      if (sh.pre_kept[ii] == 0) continue;
      PtItem it{new_off[ii], sh.pre_kept[ii], (int)pcs.size(), 0};
      pcs.insert(pcs.end(), per[ii].begin(), per[ii].end());
      it.c_end = (int)pcs.size();
      pit.push_back(it);
    }
    const int rnext = rcur ^ 1;
    dred[rnext].ensure(std::max<long long>(1, ro) * sizeof(double));
    NRGK_HIP(hipMemsetAsync(dred[rnext].p, 0, std::max<long long>(1, ro) * sizeof(double), st));
    const size_t b0 = pit.size() * sizeof(PtItem), o1 = (b0 + 255) & ~size_t(255);
    dpt.ensure(o1 + pcs.size() * sizeof(PtContrib) + 256);
    h2d((PtItem*)dpt.p, pit.data(), pit.size(), st);
    h2d((PtContrib*)(dpt.as<char>() + o1), pcs.data(), pcs.size(), st);
    launch_partial_trace(dpt.p, (int)pit.size(), dpt.as<char>() + o1, drho.as<double>(), dred[rnext].as<double>(), st);
    NRGK_HIP(stream_sync(st));
    red_off = new_off;
    rcur = rnext;
    if (ns < cap) {
      if (traces) traces[ns] = trace;
      if (sums) sums[ns] = ssum;
    }
    ns++;
  }
  d2h(pos, dpos.as<double>(), (size_t)nops * nbins, st);
  d2h(neg, dneg.as<double>(), (size_t)nops * nbins, st);
  signed_tot.assign(nops, 0.0);
  d2h(signed_tot.data(), dsum.as<double>() + 1, nops, st);
  NRGK_HIP(stream_sync(st));
  if (n_shells) *n_shells = ns;
}

// fdmThermodynamics (fdmThermodynamics.hpp:42-236) over the recorded shells.
void Fdm::thermo(double lambda, int scale_offset, int first_cnt, const double* T, int nT, double* S_out, double* C_out,
                 double* ops_out) {
  std::vector<std::pair<Shell*, double>> use;
  for (auto it = shells.rbegin(); it != shells.rend(); ++it)
    if ((*it)->cnt >= first_cnt) use.push_back({it->get(), std::pow(lambda, -((double)(*it)->cnt + scale_offset) / 2.0)});
  thermo_shells(use, loaded_nrg_max > 0 ? loaded_nrg_max : ch->cnt, T, nT, S_out, C_out, ops_out);
}

// calcThermodynamics(scale) over the given shells in order (the last shell first) and saveFinalData
// (fdmThermodynamics.hpp:42-236); nrg_max: nrg_iterations_cnt when the reference's fdmThermodynamics
// was constructed (nrgMaxIterations, the shell prefactor's exponent base, :26-29, :162)
void Fdm::thermo_shells(const std::vector<std::pair<Shell*, double>>& list, int nrg_max, const double* T, int nT,
                        double* S_out, double* C_out, double* ops_out) {
  Chain& c = *ch;
  hipStream_t st = c.st;
  std::vector<Shell*> use;
  for (auto& u : list) use.push_back(u.first);
  const int ns = (int)use.size();
  NRGK_REQUIRE(ns > 0, NRGK_ERR_STATE, "nrgk: no recorded shells");
  size_t bathDim = 0;
  for (int b = 0; b < c.bath.S; b++) bathDim += c.bath.dim(b);
  // calcThermodynamics: discarded (index >= kept of this shell) and all energies times the scale
  const int nq = ops_out ? (sysop ? sysop->nops : 0) : 0;
  std::vector<double> disc;
  std::vector<std::vector<double>> qd(nq);  // O_ll of the discarded states, aligned with disc
  std::vector<int> off{0};
  std::vector<double> gs(ns);
  for (int m = 0; m < ns; m++) {
    Shell& sh = *use[m];
    NRGK_REQUIRE(sh.q_ops >= nq && (nq == 0 || !sh.Q.empty()), NRGK_ERR_STATE,
                 "nrgk: a shell was recorded without the system operator");
    const double scale = list[m].second;
    const std::vector<int> kc = m == 0 ? std::vector<int>(sh.n.size(), 0) : use[m - 1]->pre_kept;
    NRGK_REQUIRE(kc.size() == sh.n.size(), NRGK_ERR_STATE, "nrgk: recorded shells are not consecutive");
    double mn = 1e308;
    for (size_t i = 0; i < sh.n.size(); i++)
      for (int j = 0; j < sh.n[i]; j++) mn = std::min(mn, sh.E[sh.eoff[i] + j] * scale);
    for (size_t i = 0; i < sh.n.size(); i++)
      for (int j = kc[i]; j < sh.n[i]; j++) {
        disc.push_back(sh.E[sh.eoff[i] + j] * scale - mn);
        for (int q = 0; q < nq; q++) qd[q].push_back(sh.Q[(size_t)q * sh.E.size() + sh.eoff[i] + j]);
      }
    off.push_back((int)disc.size());
    gs[m] = mn;
  }
  // saveFinalData: pairwise ground-energy correction, then relative to the last shell (:133-142)
  for (int m = 0; m + 1 < ns; m++) gs[m] += gs[m + 1];
  const double g0 = gs[0];
  for (auto& x : gs) x = x - g0;
  dE.ensure(std::max<size_t>(1, disc.size()) * sizeof(double));
  const size_t nz = (size_t)nT * ns;
  dth.ensure((3 * nz + ns + 1 + nT) * sizeof(double) + off.size() * sizeof(int) + 256);
  double* dz = dth.as<double>();
  double* dm1 = dz + nz;
  double* dm2 = dm1 + nz;
  double* dgs = dm2 + nz;
  double* dT = dgs + ns;
  int* doff = (int*)(dT + nT);
  h2d(dE.as<double>(), disc.data(), disc.size(), st);
  h2d(dgs, gs.data(), ns, st);
  h2d(dT, T, nT, st);
  h2d(doff, off.data(), off.size(), st);
  launch_thermo(dE.as<double>(), doff, dgs, ns, dT, nT, dz, dm1, dm2, st);
  std::vector<double> z(nz), m1(nz), m2(nz), mq((size_t)nq * nz);
  if (nq) {
    const size_t nd = std::max<size_t>(1, disc.size());
    dQ.ensure(((size_t)nq * nd + (size_t)nq * nz) * sizeof(double));
    double* dq = dQ.as<double>();
    double* dmq = dq + (size_t)nq * nd;
    for (int q = 0; q < nq; q++) h2d(dq + (size_t)q * nd, qd[q].data(), qd[q].size(), st);
    launch_thermo_ops(dE.as<double>(), dq, (long long)nd, nq, doff, dgs, ns, dT, nT, dmq, st);
    d2h(mq.data(), dmq, mq.size(), st);
  }
  d2h(z.data(), dz, nz, st);
  d2h(m1.data(), dm1, nz, st);
  d2h(m2.data(), dm2, nz, st);
  NRGK_HIP(stream_sync(st));
  const int nrgMax = nrg_max;
  for (int t = 0; t < nT; t++) {
    const double kT = T[t];
    std::vector<double> wm(ns);
    int cnt = nrgMax;
    for (int m = 0; m < ns; m++) {
      const double prefac = std::pow((double)bathDim, nrgMax - cnt);
      cnt--;
      wm[m] = z[(size_t)t * ns + m] * prefac;
    }
    double lpart = 0;
    for (double w : wm) lpart += w;
    for (auto& w : wm) w = w / lpart;
    double eAv = 0, eSq = 0;
    for (int m = 0; m < ns; m++) {
      const double zp = z[(size_t)t * ns + m];
      if (zp == 0.0) continue;
      eAv += wm[m] * m1[(size_t)t * ns + m] / zp;
      eSq += wm[m] * m2[(size_t)t * ns + m] / zp;
    }
    S_out[t] = std::log(lpart) + eAv / kT;
    C_out[t] = (eSq - eAv * eAv) / (kT * kT);
    // qsysOPeratorExpValue[iq][itm] (fdmThermodynamics.hpp:191-206)
    for (int q = 0; q < nq; q++) {
      double v = 0;
      for (int m = 0; m < ns; m++) {
        const double zp = z[(size_t)t * ns + m];
        if (zp == 0.0) continue;
        v += wm[m] * mq[((size_t)t * nq + q) * ns + m] / zp;
      }
      ops_out[(size_t)q * nT + t] = v;
    }
  }
}

}  // namespace nrgk
