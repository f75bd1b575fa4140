// extern "C" entry points of libnrgk (declared in include/nrgk.h).  Every call converts exceptions
// into a status code and keeps the message for nrgk_last_error().
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <system_error>
#include <thread>

#include "chain.hpp"
#include "h5io.hpp"
#include "kernels.hpp"
#include "observables.hpp"
#include "server.hpp"

namespace nrgk {
static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }

template <class F>
static int guard(F&& f) {
  try {
    f();
    return NRGK_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.status;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return NRGK_ERR_INVALID;
  }
}
}  // namespace nrgk

using namespace nrgk;

struct nrgk_ctx {
  Ctx c;
};
struct nrgk_chain {
  std::unique_ptr<Chain> c;
};
struct nrgk_server {
  std::unique_ptr<Server> s;
};

extern "C" {

const char* nrgk_last_error(void) { return g_last_error.c_str(); }
const char* nrgk_version(void) { return "nrgk 0.1 (gfx950)"; }

int nrgk_ctx_create(int device, nrgk_ctx** out) {
  return guard([&] {
    NRGK_REQUIRE(out, NRGK_ERR_INVALID, "nrgk_ctx_create: null output");
    int ndev = 0;
    NRGK_HIP(hipGetDeviceCount(&ndev));
    NRGK_REQUIRE(device >= 0 && device < ndev, NRGK_ERR_INVALID, "nrgk_ctx_create: no such device");
    NRGK_HIP(hipSetDevice(device));
    auto* h = new nrgk_ctx;
    h->c.device = device;
    hipDeviceProp_t prop;
    NRGK_HIP(hipGetDeviceProperties(&prop, device));
    h->c.n_cu = prop.multiProcessorCount;
    int lds = (int)prop.sharedMemPerBlock;
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, device) == hipSuccess && optin > lds)
      lds = optin;
    h->c.lds_max = lds - 1024;  // room for the kernels' static LDS
    h->c.arch = prop.gcnArchName;
    const size_t colon = h->c.arch.find(':');
    if (colon != std::string::npos) h->c.arch.resize(colon);
    set_kernel_lds_limits(h->c.lds_max);
    set_ht_lds_limits(h->c.lds_max);
    noalloc_scratch_setup(device);
    tune_load(h->c);
    size_t fr = 0, tot = 0;
    NRGK_HIP(hipMemGetInfo(&fr, &tot));
    h->c.capacity = (size_t)(0.9 * (double)fr);
    *out = h;
  });
}
int nrgk_tune(nrgk_ctx* ctx, const char* cache_path, int* crossover, double* table, int cap_rows, int* n_rows) {
  return guard([&] {
    NRGK_REQUIRE(ctx, NRGK_ERR_INVALID, "null context");
    NRGK_HIP(hipSetDevice(ctx->c.device));
    std::vector<std::vector<double>> t;
    const std::string path = cache_path ? std::string(cache_path) : tune_cache_path(ctx->c);
    const int c = tune_crossover(&ctx->c, path, t);
    if (crossover) *crossover = c;
    if (n_rows) *n_rows = (int)t.size();
    if (table)
      for (int r = 0; r < std::min(cap_rows, (int)t.size()); r++)
        for (int k = 0; k < 4; k++) table[4 * r + k] = t[r][k];
  });
}
int nrgk_ctx_crossover(nrgk_ctx* ctx, int* crossover, int* tuned) {
  return guard([&] {
    NRGK_REQUIRE(ctx, NRGK_ERR_INVALID, "null context");
    if (crossover) *crossover = ctx->c.mixed_max;
    if (tuned) *tuned = ctx->c.tuned ? 1 : 0;
  });
}
int nrgk_ctx_destroy(nrgk_ctx* ctx) {
  delete ctx;
  return NRGK_OK;
}
int nrgk_ctx_set_capacity(nrgk_ctx* ctx, size_t bytes) {
  return guard([&] {
    NRGK_REQUIRE(ctx, NRGK_ERR_INVALID, "null context");
    if (bytes == 0) {
      size_t fr = 0, tot = 0;
      NRGK_HIP(hipMemGetInfo(&fr, &tot));
      bytes = (size_t)(0.9 * (double)fr) + ctx->c.used;
    }
    ctx->c.capacity = bytes;
  });
}
int nrgk_ctx_used_bytes(nrgk_ctx* ctx, size_t* used, size_t* capacity) {
  return guard([&] {
    NRGK_REQUIRE(ctx, NRGK_ERR_INVALID, "null context");
    if (used) *used = ctx->c.used;
    if (capacity) *capacity = ctx->c.capacity;
  });
}

int nrgk_chain_create(nrgk_ctx* ctx, const nrgk_site_desc* impurity, const nrgk_site_desc* bath, size_t max_kept,
                      int mode, nrgk_chain** out) {
  return guard([&] {
    NRGK_REQUIRE(ctx && out, NRGK_ERR_INVALID, "nrgk_chain_create: null argument");
    auto* h = new nrgk_chain;
    try {
      h->c = std::make_unique<Chain>(&ctx->c, impurity, bath, max_kept, mode);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int nrgk_chain_destroy(nrgk_chain* chain) {
  delete chain;
  return NRGK_OK;
}
int nrgk_chain_set_parameters(nrgk_chain* ch, size_t max_kept) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    Chain& c = *ch->c;
    c.K = max_kept;
    c.no_kept = max_kept;
    c.errbar = 1e-8;
    c.cnt = 0;
    c.cnt_min = 0;
    c.phase = Chain::FRESH;
  });
}
int nrgk_chain_set_bath(nrgk_chain* ch, const nrgk_site_desc* bath) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    ch->c->set_bath(bath);
  });
}
int nrgk_chain_add_bath_site(nrgk_chain* ch, const double* t, int n, double rescale) {
  return guard([&] {
    NRGK_REQUIRE(ch && (t || n == 0), NRGK_ERR_INVALID, "null argument");
    ch->c->add_bath_site(t, n, rescale);
  });
}
int nrgk_chain_update_internal_state(nrgk_chain* ch) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    ch->c->update_internal_state();
  });
}
int nrgk_chain_step(nrgk_chain* ch, const double* t, int n, double rescale) {
  return guard([&] {
    NRGK_REQUIRE(ch && (t || n == 0), NRGK_ERR_INVALID, "null argument");
    ch->c->add_bath_site(t, n, rescale);
    ch->c->update_internal_state();
  });
}
int nrgk_chain_counters(nrgk_chain* ch, int* cnt, int* cnt_min) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    if (cnt) *cnt = ch->c->cnt;
    if (cnt_min) *cnt_min = ch->c->cnt_min;
  });
}
int nrgk_chain_set_counter(nrgk_chain* ch, int cnt) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    ch->c->cnt = cnt;
  });
}
int nrgk_chain_higher_energy_discarded(nrgk_chain* ch, int* out) {
  return guard([&] {
    NRGK_REQUIRE(ch && out, NRGK_ERR_INVALID, "null argument");
    *out = ch->c->no_kept >= ch->c->K ? 1 : 0;  // checkHigherEnergyDiscarded (nrgcore.hpp:384-386)
  });
}
int nrgk_chain_all_eigenvalues(nrgk_chain* ch, double* out, size_t cap, size_t* n) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    Chain& c = *ch->c;
    NRGK_REQUIRE(c.phase != Chain::FRESH || !c.all_eig.empty(), NRGK_ERR_STATE, "nrgk: no eigenvalues yet");
    c.ensure_all_eig();
    if (n) *n = c.all_eig.size();
    if (out) {
      NRGK_REQUIRE(cap >= c.all_eig.size(), NRGK_ERR_INVALID, "nrgk: output capacity too small");
      std::memcpy(out, c.all_eig.data(), c.all_eig.size() * sizeof(double));
    }
  });
}
int nrgk_chain_relative_gs(nrgk_chain* ch, double* out, size_t cap, size_t* n) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    auto& v = ch->c->relgs;
    if (n) *n = v.size();
    if (out) {
      NRGK_REQUIRE(cap >= v.size(), NRGK_ERR_INVALID, "nrgk: output capacity too small");
      std::memcpy(out, v.data(), v.size() * sizeof(double));
    }
  });
}
int nrgk_chain_sectors(nrgk_chain* ch, int* n_sectors, int* qn, int* dims) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    Chain& c = *ch->c;
    NRGK_REQUIRE(!c.cur_qn.empty(), NRGK_ERR_STATE, "current_sysmQ.empty!");
    if (n_sectors) *n_sectors = (int)c.cur_qn.size();
    for (size_t s = 0; s < c.cur_qn.size(); s++) {
      if (qn)
        for (int q = 0; q < c.imp.nqn; q++) qn[s * c.imp.nqn + q] = c.cur_qn[s][q];
      if (dims) dims[s] = c.cur_n[s];
    }
  });
}
int nrgk_chain_pre_sectors(nrgk_chain* ch, int* n_sectors, int* qn) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    Chain& c = *ch->c;
    if (n_sectors) *n_sectors = (int)c.pre_qn.size();
    if (qn)
      for (size_t s = 0; s < c.pre_qn.size(); s++)
        for (int q = 0; q < c.imp.nqn; q++) qn[s * c.imp.nqn + q] = c.pre_qn[s][q];
  });
}
int nrgk_chain_coupled_index(nrgk_chain* ch, int* counts, int* flat, size_t cap, size_t* total) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    Chain& c = *ch->c;
    size_t t = 0;
    for (auto& v : c.coupled) t += v.size();
    if (total) *total = t;
    if (counts)
      for (size_t s = 0; s < c.coupled.size(); s++) counts[s] = (int)c.coupled[s].size();
    if (flat) {
      NRGK_REQUIRE(cap >= t, NRGK_ERR_INVALID, "nrgk: output capacity too small");
      size_t o = 0;
      for (auto& v : c.coupled)
        for (auto [ii, bb] : v) flat[o++] = ii * c.bath.S + bb;  // coupled_nQ_index (nrgcore.hpp:238-261)
    }
  });
}
int nrgk_chain_eigenvalues_q(nrgk_chain* ch, double* out, size_t cap, size_t* n) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    Chain& c = *ch->c;
    NRGK_REQUIRE(c.phase != Chain::FRESH, NRGK_ERR_STATE, "eigenvaluesQ.empty!");
    if (n) *n = (size_t)c.N_tot;
    if (out) {
      NRGK_REQUIRE(cap >= (size_t)c.N_tot, NRGK_ERR_INVALID, "nrgk: output capacity too small");
      const double* src = c.phase == Chain::ASSEMBLED ? c.dlam_raw.as<double>() : c.dlam_shift[c.lcur].as<double>();
      d2h(out, src, c.N_tot, c.st);
      NRGK_HIP(stream_sync(c.st));
    }
  });
}
int nrgk_chain_kept_counts(nrgk_chain* ch, int* out, int cap, int* n_sectors) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    Chain& c = *ch->c;
    const std::vector<int>& k = c.phase == Chain::TRUNCATED ? c.kept : c.pre_kept;
    if (n_sectors) *n_sectors = (int)k.size();
    if (out) {
      NRGK_REQUIRE(cap >= (int)k.size(), NRGK_ERR_INVALID, "nrgk: output capacity too small");
      std::copy(k.begin(), k.end(), out);
    }
  });
}
int nrgk_chain_block_u(nrgk_chain* ch, int s, double* out, size_t cap) {
  return guard([&] {
    NRGK_REQUIRE(ch && out, NRGK_ERR_INVALID, "null argument");
    Chain& c = *ch->c;
    NRGK_REQUIRE(c.phase != Chain::FRESH && s >= 0 && s < (int)c.cur_n.size(), NRGK_ERR_INVALID, "nrgk: bad sector");
    const size_t sz = (size_t)c.cur_n[s] * c.cur_n[s];
    NRGK_REQUIRE(cap >= sz, NRGK_ERR_INVALID, "nrgk: output capacity too small");
    if (c.is_wide(s)) c.wide_vectors(true);  // all columns of the wide blocks, when first asked for
    d2h(out, c.dU.as<double>() + c.cur_off[s], sz, c.st);
    NRGK_HIP(stream_sync(c.st));
  });
}
int nrgk_chain_keep_hamiltonians(nrgk_chain* ch, int on) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    ch->c->keep_h = on != 0;
  });
}
int nrgk_chain_block_h(nrgk_chain* ch, int s, double* out, size_t cap) {
  return guard([&] {
    NRGK_REQUIRE(ch && out, NRGK_ERR_INVALID, "null argument");
    Chain& c = *ch->c;
    NRGK_REQUIRE(s >= 0 && s < (int)c.h_keep.size(), NRGK_ERR_STATE, "nrgk: no kept Hamiltonian (keep_hamiltonians)");
    NRGK_REQUIRE(cap >= c.h_keep[s].size(), NRGK_ERR_INVALID, "nrgk: output capacity too small");
    std::memcpy(out, c.h_keep[s].data(), c.h_keep[s].size() * sizeof(double));
  });
}
int nrgk_chain_fdag_blocks(nrgk_chain* ch, int op, int* n_blocks, int* bi, int* bj, int* rows, int* cols,
                           double* data, size_t cap) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    Chain& c = *ch->c;
    NRGK_REQUIRE(op >= 0 && op < (int)c.fprev.size(), NRGK_ERR_INVALID, "nrgk: bad operator index");
    const OpMap& m = c.fprev[op];
    if (n_blocks) *n_blocks = (int)m.size();
    size_t need = 0;
    int b = 0;
    for (auto& [key, ref] : m) {
      if (bi) bi[b] = key.first;
      if (bj) bj[b] = key.second;
      if (rows) rows[b] = ref.rows;
      if (cols) cols[b] = ref.cols;
      need += (size_t)ref.rows * ref.cols;
      b++;
    }
    if (data) {
      NRGK_REQUIRE(cap >= need, NRGK_ERR_INVALID, "nrgk: output capacity too small");
      size_t o = 0;
      for (auto& [key, ref] : m) {
        d2h(data + o, c.fbuf[c.fcur].as<double>() + ref.off, (size_t)ref.rows * ref.cols, c.st);
        o += (size_t)ref.rows * ref.cols;
      }
      NRGK_HIP(stream_sync(c.st));
    }
  });
}
int nrgk_chain_set_state(nrgk_chain* ch, const nrgk_site_desc* state, int cnt) {
  return guard([&] {
    NRGK_REQUIRE(ch && state, NRGK_ERR_INVALID, "null argument");
    ch->c->set_state(state, cnt);
  });
}
// ---------------------------------------------------------------- observables / FDM / thermo
struct nrgk_obs {
  std::unique_ptr<Obs> o;
};
struct nrgk_fdm {
  std::unique_ptr<Fdm> f;
};

int nrgk_obs_create(nrgk_chain* ch, int n_ops, const int* op_nblocks, const int* op_bi, const int* op_bj,
                    const double* op_data, nrgk_obs** out) {
  return guard([&] {
    NRGK_REQUIRE(ch && out && n_ops > 0 && op_nblocks, NRGK_ERR_INVALID, "nrgk_obs_create: null argument");
    auto* h = new nrgk_obs;
    try {
      h->o = std::make_unique<Obs>(ch->c.get(), n_ops, op_nblocks, op_bi, op_bj, op_data);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int nrgk_obs_destroy(nrgk_obs* obs) {
  delete obs;
  return NRGK_OK;
}
int nrgk_obs_update(nrgk_obs* obs) {
  return guard([&] {
    NRGK_REQUIRE(obs, NRGK_ERR_INVALID, "null observable");
    obs->o->update();
  });
}
int nrgk_obs_blocks(nrgk_obs* obs, int op, int* n_blocks, int* bi, int* bj, int* rows, int* cols, double* data,
                    size_t cap) {
  return guard([&] {
    NRGK_REQUIRE(obs && op >= 0 && op < obs->o->nops, NRGK_ERR_INVALID, "nrgk: bad observable index");
    Obs& o = *obs->o;
    const OpMap& m = o.blocks[op];
    if (n_blocks) *n_blocks = (int)m.size();
    size_t need = 0;
    int b = 0;
    for (auto& [key, ref] : m) {
      if (bi) bi[b] = key.first;
      if (bj) bj[b] = key.second;
      if (rows) rows[b] = ref.rows;
      if (cols) cols[b] = ref.cols;
      need += (size_t)ref.rows * ref.cols;
      b++;
    }
    if (data) {
      NRGK_REQUIRE(cap >= need, NRGK_ERR_INVALID, "nrgk: output capacity too small");
      size_t off = 0;
      for (auto& [key, ref] : m) {
        d2h(data + off, o.buf[o.cur].as<double>() + ref.off, (size_t)ref.rows * ref.cols, o.ch->st);
        off += (size_t)ref.rows * ref.cols;
      }
      NRGK_HIP(stream_sync(o.ch->st));
    }
  });
}
int nrgk_fdm_create(nrgk_chain* ch, nrgk_obs* obs, int n_bins, double e_min, double e_max, nrgk_fdm** out) {
  return guard([&] {
    NRGK_REQUIRE(ch && out, NRGK_ERR_INVALID, "nrgk_fdm_create: null argument");
    auto* h = new nrgk_fdm;
    try {
      h->f = std::make_unique<Fdm>(ch->c.get(), obs ? obs->o.get() : nullptr, n_bins, e_min, e_max);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int nrgk_fdm_destroy(nrgk_fdm* fdm) {
  delete fdm;
  return NRGK_OK;
}
int nrgk_fdm_record(nrgk_fdm* fdm) {
  return guard([&] {
    NRGK_REQUIRE(fdm, NRGK_ERR_INVALID, "null recorder");
    fdm->f->record();
  });
}
int nrgk_fdm_backward(nrgk_fdm* fdm, double lambda, int scale_offset, int first_cnt, double* pos, double* neg,
                      double* traces, double* sums, int cap_shells, int* n_shells) {
  return guard([&] {
    NRGK_REQUIRE(fdm && pos && neg, NRGK_ERR_INVALID, "nrgk_fdm_backward: null argument");
    fdm->f->backward(lambda, scale_offset, first_cnt, pos, neg, traces, sums, cap_shells, n_shells);
  });
}
int nrgk_fdm_energy_points(nrgk_fdm* fdm, double* out) {
  return guard([&] {
    NRGK_REQUIRE(fdm && out, NRGK_ERR_INVALID, "null argument");
    const Fdm& f = *fdm->f;
    for (int i = 0; i < f.nbins; i++) out[i] = f.e_min * std::exp(i * 1. / f.delE);
  });
}
int nrgk_fdm_add_shell(nrgk_fdm* fdm, int cnt, double scale) {
  return guard([&] {
    NRGK_REQUIRE(fdm && std::isfinite(scale), NRGK_ERR_INVALID, "nrgk_fdm_add_shell: bad argument");
    fdm->f->shell_at(cnt);  // must exist
    fdm->f->registered.push_back({cnt, scale});
  });
}
int nrgk_fdm_run_spectrum(nrgk_fdm* fdm, double* pos, double* neg, double* traces, double* sums, int cap_shells,
                          int* n_shells) {
  return guard([&] {
    NRGK_REQUIRE(fdm && pos && neg, NRGK_ERR_INVALID, "nrgk_fdm_run_spectrum: null argument");
    Fdm& f = *fdm->f;
    NRGK_REQUIRE(!f.registered.empty(), NRGK_ERR_STATE, "nrgk_fdm_run_spectrum: no shell registered");
    std::vector<std::pair<Shell*, double>> use;
    for (auto& [cnt, sc] : f.registered) use.push_back({f.shell_at(cnt), sc});
    f.registered.clear();
    f.backward_shells(use, pos, neg, traces, sums, cap_shells, n_shells);
  });
}
int nrgk_thermo_run(nrgk_fdm* fdm, int nrg_max, const double* T, int nT, double* S, double* Cv, double* ops) {
  return guard([&] {
    NRGK_REQUIRE(fdm && T && S && Cv && nT > 0, NRGK_ERR_INVALID, "nrgk_thermo_run: null argument");
    Fdm& f = *fdm->f;
    NRGK_REQUIRE(!f.registered.empty(), NRGK_ERR_STATE, "nrgk_thermo_run: no shell registered");
    std::vector<std::pair<Shell*, double>> use;
    for (auto& [cnt, sc] : f.registered) use.push_back({f.shell_at(cnt), sc});
    f.registered.clear();
    f.thermo_shells(use, nrg_max, T, nT, S, Cv, ops);
  });
}
int nrgk_h5_create(const char* path) {
  return guard([&] {
    NRGK_REQUIRE(path, NRGK_ERR_INVALID, "nrgk_h5_create: null path");
    h5::File f(path, true);
  });
}
int nrgk_h5_write(const char* path, const char* name, const double* data, size_t n) {
  return guard([&] {
    NRGK_REQUIRE(path && name && (data || n == 0), NRGK_ERR_INVALID, "nrgk_h5_write: null argument");
    h5::File f(path, h5::File::Mode::ReadWrite);
    const std::string nm(name);
    const size_t sl = nm.find_last_of('/');
    if (sl != std::string::npos && sl > 0) f.group(nm.substr(0, sl));
    f.write(nm, h5::Type::F64, data, n);
  });
}
int nrgk_fdm_signed_weights(nrgk_fdm* fdm, double* out, int cap, int* n_ops) {
  return guard([&] {
    NRGK_REQUIRE(fdm, NRGK_ERR_INVALID, "null argument");
    const Fdm& f = *fdm->f;
    NRGK_REQUIRE(!f.signed_tot.empty(), NRGK_ERR_STATE, "nrgk_fdm_signed_weights: no backward pass yet");
    if (n_ops) *n_ops = (int)f.signed_tot.size();
    if (out)
      for (int p = 0; p < std::min<int>(cap, (int)f.signed_tot.size()); p++) out[p] = f.signed_tot[p];
  });
}
int nrgk_thermo(nrgk_fdm* rec, double lambda, int scale_offset, int first_cnt, const double* T, int nT, double* S,
                double* C) {
  return guard([&] {
    NRGK_REQUIRE(rec && T && S && C && nT > 0, NRGK_ERR_INVALID, "nrgk_thermo: null argument");
    rec->f->thermo(lambda, scale_offset, first_cnt, T, nT, S, C);
  });
}
int nrgk_fdm_load(nrgk_fdm* rec, const char* path, const char* op_group) {
  return guard([&] {
    NRGK_REQUIRE(rec && path, NRGK_ERR_INVALID, "nrgk_fdm_load: null argument");
    rec->f->load(path, op_group ? op_group : "");
  });
}
int nrgk_fdm_save(nrgk_fdm* rec, const char* path, const char* op_group) {
  return guard([&] {
    NRGK_REQUIRE(rec && path, NRGK_ERR_INVALID, "nrgk_fdm_save: null argument");
    rec->f->save(path, op_group ? op_group : "");
  });
}
int nrgk_h5_read(const char* path, const char* dataset, double* out, size_t cap, size_t* n) {
  return guard([&] {
    NRGK_REQUIRE(path && dataset, NRGK_ERR_INVALID, "nrgk_h5_read: null argument");
    h5::File f(path, false);
    const size_t m = f.size(dataset);
    if (n) *n = m;
    if (!out) return;
    NRGK_REQUIRE(cap >= m, NRGK_ERR_INVALID, "nrgk_h5_read: output capacity too small");
    f.read(dataset, h5::Type::F64, out, m);
  });
}
int nrgk_fdm_set_system_operator(nrgk_fdm* rec, nrgk_obs* sysop) {
  return guard([&] {
    NRGK_REQUIRE(rec && sysop, NRGK_ERR_INVALID, "nrgk_fdm_set_system_operator: null argument");
    NRGK_REQUIRE(sysop->o->ch == rec->f->ch, NRGK_ERR_INVALID,
                 "nrgk_fdm_set_system_operator: the operator belongs to another chain");
    rec->f->sysop = sysop->o.get();
  });
}
int nrgk_thermo_expectation(nrgk_fdm* rec, double lambda, int scale_offset, int first_cnt, const double* T, int nT,
                            double* out) {
  return guard([&] {
    NRGK_REQUIRE(rec && T && out && nT > 0, NRGK_ERR_INVALID, "nrgk_thermo_expectation: null argument");
    NRGK_REQUIRE(rec->f->sysop, NRGK_ERR_STATE, "nrgk_thermo_expectation: no system operator set");
    std::vector<double> S(nT), C(nT);
    rec->f->thermo(lambda, scale_offset, first_cnt, T, nT, S.data(), C.data(), out);
  });
}
int nrgk_server_create(nrgk_ctx* ctx, int n_slots, size_t hbm_budget, nrgk_server** out) {
  return guard([&] {
    NRGK_REQUIRE(ctx && out, NRGK_ERR_INVALID, "nrgk_server_create: null argument");
    auto* h = new nrgk_server;
    try {
      h->s = std::make_unique<Server>(&ctx->c, n_slots, hbm_budget);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int nrgk_server_destroy(nrgk_server* srv) {
  return guard([&] {
    if (!srv) return;
    {
      std::unique_lock<std::mutex> lk(srv->s->m);
      srv->s->cv.wait(lk, [&] { return srv->s->pending.empty() && srv->s->running == 0; });
    }
    delete srv;
  });
}
int nrgk_server_submit(nrgk_server* srv, const nrgk_request* q, long long* id) {
  return guard([&] {
    NRGK_REQUIRE(srv && q && id && q->impurity && q->bath && q->observable, NRGK_ERR_INVALID,
                 "nrgk_server_submit: null argument");
    NRGK_REQUIRE(q->n_omega == 0 || q->omega, NRGK_ERR_INVALID, "nrgk_server_submit: null omega");
    Request r;
    r.impurity = OwnedSite::copy(q->impurity);
    r.bath = OwnedSite::copy(q->bath);
    r.observable = OwnedSite::copy(q->observable);
    NRGK_REQUIRE(r.observable.n_ops > 0, NRGK_ERR_INVALID, "nrgk_server_submit: the observable has no operator");
    r.max_kept = q->max_kept;
    r.lambda = q->lambda;
    r.first_hopping = q->first_hopping;
    r.n_sites = q->n_sites;
    r.n_bins = q->n_bins;
    r.n_omega = q->n_omega;
    r.omega.assign(q->omega, q->omega + std::max(0, q->n_omega));
    r.broadening = q->broadening;
    r.fail_at_site = q->fail_at_site;
    *id = srv->s->submit(std::move(r));
  });
}
int nrgk_server_poll(nrgk_server* srv, long long id, int* done) {
  return guard([&] {
    NRGK_REQUIRE(srv && done, NRGK_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(srv->s->m);
    auto it = srv->s->all.find(id);
    NRGK_REQUIRE(it != srv->s->all.end(), NRGK_ERR_INVALID, "nrgk_server_poll: unknown request");
    *done = it->second->done ? 1 : 0;
  });
}
int nrgk_server_wait(nrgk_server* srv, long long id, double* A_pos, double* A_neg, double* info) {
  std::shared_ptr<Request> r;
  const int rc = guard([&] {
    NRGK_REQUIRE(srv, NRGK_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(srv->s->m);
    auto it = srv->s->all.find(id);
    NRGK_REQUIRE(it != srv->s->all.end(), NRGK_ERR_INVALID, "nrgk_server_wait: unknown request");
    r = it->second;
    srv->s->cv.wait(lk, [&] { return r->done; });
    srv->s->all.erase(it);
  });
  if (rc != NRGK_OK) return rc;
  using ms = std::chrono::duration<double, std::milli>;
  if (A_pos) std::copy(r->A_pos.begin(), r->A_pos.end(), A_pos);
  if (A_neg) std::copy(r->A_neg.begin(), r->A_neg.end(), A_neg);
  if (info) {
    info[0] = ms(r->t_done - r->t_submit).count();
    info[1] = ms(r->t_start - r->t_submit).count();
    info[2] = r->total_weight;
    info[3] = r->iterations;
    info[4] = r->shells;
    info[5] = (double)r->bytes;
  }
  if (r->status != NRGK_OK) set_last_error(r->message);
  return r->status;
}
int nrgk_server_stats(nrgk_server* srv, int* running, int* queued, size_t* reserved, int* peak_running,
                      size_t* peak_reserved) {
  return guard([&] {
    NRGK_REQUIRE(srv, NRGK_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(srv->s->m);
    if (running) *running = srv->s->running;
    if (queued) *queued = (int)srv->s->pending.size();
    if (reserved) *reserved = srv->s->reserved;
    if (peak_running) *peak_running = srv->s->peak_running;
    if (peak_reserved) *peak_reserved = srv->s->peak_reserved;
  });
}

int nrgk_broaden(nrgk_ctx* ctx, const double* x, const double* w, int n, const double* omega, int n_omega, double b,
                 double* out) {
  return guard([&] {
    NRGK_REQUIRE(ctx && x && w && omega && out && n >= 0 && n_omega >= 0 && b > 0, NRGK_ERR_INVALID,
                 "nrgk_broaden: bad argument");
    NRGK_HIP(hipSetDevice(ctx->c.device));
    DevBuf d(&ctx->c);
    const size_t ns = broaden_scratch_doubles(n, n_omega);
    d.ensure(((size_t)2 * n + 2 * n_omega + ns + 4) * sizeof(double));
    double* dx = d.as<double>();
    double* dw = dx + n;
    double* dom = dw + n;
    double* dout = dom + n_omega;
    double* dscr = dout + n_omega;
    h2d(dx, x, n, nullptr);
    h2d(dw, w, n, nullptr);
    h2d(dom, omega, n_omega, nullptr);
    launch_broaden(dx, dw, n, dom, n_omega, b, dout, dscr, nullptr);
    d2h(out, dout, n_omega, nullptr);
    NRGK_HIP(stream_sync(nullptr));
  });
}
// nrgk_broaden's kernels on device buffers the caller owns, with the fixed scratch of noalloc.hip:
// enqueues on the caller's stream and allocates nothing (no hipMalloc, no library allocation)
int nrgk_broaden_noalloc(nrgk_ctx* ctx, const double* x, const double* w, int n, const double* omega, int n_omega,
                         double b, double* out, void* stream) {
  return guard([&] {
    NRGK_REQUIRE(ctx && x && w && omega && out && n >= 0 && n_omega >= 0 && b > 0, NRGK_ERR_INVALID,
                 "nrgk_broaden_noalloc: bad argument");
    const size_t need = broaden_scratch_doubles(n, n_omega);
    NRGK_REQUIRE(need <= NOALLOC_SCRATCH_DOUBLES, NRGK_ERR_INVALID,
                 "nrgk_broaden_noalloc: ceil(n / 512) * n_omega = " + std::to_string(need) +
                     " partial sums exceed the fixed scratch of " + std::to_string(NOALLOC_SCRATCH_DOUBLES) +
                     " doubles");
    double* scr = noalloc_scratch(ctx->c.device);
    NRGK_REQUIRE(scr, NRGK_ERR_STATE, "nrgk_broaden_noalloc: the context's device has no scratch address");
    NRGK_HIP(hipSetDevice(ctx->c.device));
    launch_broaden(x, w, n, omega, n_omega, b, out, scr, static_cast<hipStream_t>(stream));
  });
}

int nrgk_chain_refine_stats(nrgk_chain* ch, int* oa_steps, int* unconverged) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    if (oa_steps) *oa_steps = ch->c->oa_steps;
    if (unconverged) *unconverged = ch->c->oa_unconverged;
  });
}
int nrgk_chain_profile(nrgk_chain* ch, int enable) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    ch->c->sync();
    ch->c->prof_on = enable != 0;
    if (enable) ch->c->kstats.clear();
  });
}
int nrgk_chain_profile_only(nrgk_chain* ch, const char* family) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    ch->c->prof_only = family ? family : "";
  });
}
int nrgk_chain_kernel_stats(nrgk_chain* ch, nrgk_kernel_stat* out, int cap, int* n_out) {
  return guard([&] {
    NRGK_REQUIRE(ch && n_out && (out || cap == 0), NRGK_ERR_INVALID, "null argument");
    ch->c->sync();  // runs still in flight are resolved
    int i = 0;
    for (auto& [name, k] : ch->c->kstats) {
      if (i < cap) {
        std::memset(out[i].family, 0, sizeof(out[i].family));
        std::strncpy(out[i].family, name.c_str(), sizeof(out[i].family) - 1);
        out[i].ms = k.ms;
        out[i].launches = k.launches;
        out[i].flops = k.flops;
        out[i].bytes = k.bytes;
      }
      i++;
    }
    *n_out = i;
  });
}
int nrgk_chain_step_stats(nrgk_chain* ch, double* ms4, int* sweeps, int* max_cluster) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    Chain& c = *ch->c;
    if (ms4)
      for (int i = 0; i < 4; i++) ms4[i] = c.ms[i];
    if (sweeps) *sweeps = c.sweeps;
    if (max_cluster) *max_cluster = c.unresolved ? -c.maxcluster : c.maxcluster;
  });
}

int nrgk_chain_truncation_stats(nrgk_chain* ch, double* e_cut, int* near_ties) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    if (e_cut) *e_cut = ch->c->last_cut;
    if (near_ties) *near_ties = ch->c->near_ties;
  });
}
int nrgk_chain_step_counts(nrgk_chain* ch, long long* launches, long long* syncs) {
  return guard([&] {
    NRGK_REQUIRE(ch, NRGK_ERR_INVALID, "null chain");
    if (launches) *launches = ch->c->last_launches;
    if (syncs) *syncs = ch->c->last_syncs;
  });
}

// Several independent chains advanced by one Wilson site each, concurrently.  Each chain runs on its
// own stream; one host worker per chain at a time drives it, so that the chains' latency-bound launch
// sequences (and the host waits between them) overlap on the GPU's hardware queues.  The per-chain
// order is the loop of the reference's FDM drivers (examples/fdmSpectrumSiam/main.cpp:86-95).
int nrgk_chains_step(nrgk_chain* const* chains, int n_chains, const double* t, int nt, const double* rescale,
                     nrgk_obs* const* obs, int n_obs, nrgk_fdm* const* recorders, int n_threads, int* recorded) {
  return guard([&] {
    NRGK_REQUIRE(n_chains >= 0 && nt >= 0 && n_obs >= 0, NRGK_ERR_INVALID, "nrgk_chains_step: bad sizes");
    if (n_chains == 0) return;
    NRGK_REQUIRE(chains && rescale && (t || nt == 0) && (obs || n_obs == 0), NRGK_ERR_INVALID,
                 "nrgk_chains_step: null argument");
    std::vector<const Chain*> seen;
    for (int i = 0; i < n_chains; i++) {
      NRGK_REQUIRE(chains[i] && chains[i]->c, NRGK_ERR_INVALID, "nrgk_chains_step: null chain");
      seen.push_back(chains[i]->c.get());
      for (int k = 0; k < n_obs; k++) {
        nrgk_obs* o = obs[(size_t)i * n_obs + k];
        NRGK_REQUIRE(o && o->o->ch == chains[i]->c.get(), NRGK_ERR_INVALID,
                     "nrgk_chains_step: an observable does not belong to its chain");
      }
      if (recorders && recorders[i])
        NRGK_REQUIRE(recorders[i]->f->ch == chains[i]->c.get(), NRGK_ERR_INVALID,
                     "nrgk_chains_step: a recorder does not belong to its chain");
    }
    std::sort(seen.begin(), seen.end());
    NRGK_REQUIRE(std::adjacent_find(seen.begin(), seen.end()) == seen.end(), NRGK_ERR_INVALID,
                 "nrgk_chains_step: a chain appears twice");
    std::vector<int> status(n_chains, NRGK_OK), rec(n_chains, 0);
    std::vector<std::string> msg(n_chains);
    std::atomic<int> next{0};
    auto work = [&] {
      for (int i; (i = next.fetch_add(1)) < n_chains;) {
        Chain& c = *chains[i]->c;
        try {
          NRGK_HIP(hipSetDevice(c.ctx->device));
          c.add_bath_site(t + (size_t)i * nt, nt, rescale[i]);
          for (int k = 0; k < n_obs; k++) obs[(size_t)i * n_obs + k]->o->update();
          if (recorders && recorders[i] && c.no_kept >= c.K) {  // checkHigherEnergyDiscarded
            recorders[i]->f->record();
            rec[i] = 1;
          }
          c.update_internal_state();
        } catch (const Error& e) {
          status[i] = e.status;
          msg[i] = e.what();
        } catch (const std::exception& e) {
          status[i] = NRGK_ERR_INVALID;
          msg[i] = e.what();
        }
      }
    };
    const int nw = std::max(1, std::min(n_chains, n_threads > 0 ? n_threads : 16));
    // NRGK_DEBUG_MAX_WORKERS=k: thread creation fails after k - 1 workers (the path below, on demand)
    const char* mw = std::getenv("NRGK_DEBUG_MAX_WORKERS");
    const int max_started = mw ? std::max(0, std::atoi(mw) - 1) : nw;
    std::vector<std::thread> pool;
    try {
      for (int w = 1; w < nw; w++) {
        if ((int)pool.size() >= max_started)
          throw std::system_error(std::make_error_code(std::errc::resource_unavailable_try_again));
        pool.emplace_back(work);
      }
    } catch (const std::system_error&) {
      // no more threads: the workers that started and this thread take the remaining chains
    }
    work();
    for (auto& th : pool) th.join();
    if (recorded) std::copy(rec.begin(), rec.end(), recorded);
    for (int i = 0; i < n_chains; i++)
      if (status[i] != NRGK_OK) throw Error(status[i], "chain " + std::to_string(i) + ": " + msg[i]);
  });
}

// qmatrix::diag (qmatrix.hpp:375-404) for a list of blocks in one call, through the chain's eigensolver.
int nrgk_eigh_batched(nrgk_ctx* ctx, int n_blocks, const int* dims, const double* H, double* w, double* U,
                      int mode) {
  return guard([&] {
    NRGK_REQUIRE(ctx && n_blocks >= 0 && (dims || n_blocks == 0), NRGK_ERR_INVALID, "nrgk_eigh_batched: bad argument");
    std::vector<int> d(dims, dims + n_blocks);
    long long ne = 0;
    for (int x : d) {
      NRGK_REQUIRE(x >= 0, NRGK_ERR_INVALID, "nrgk_eigh_batched: negative block size");
      ne += x;
    }
    NRGK_REQUIRE(mode == NRGK_MODE_MIXED || mode == NRGK_MODE_PLAIN64, NRGK_ERR_INVALID, "nrgk: unknown mode");
    if (ne == 0) return;
    NRGK_REQUIRE(H && w && U, NRGK_ERR_INVALID, "nrgk_eigh_batched: null array");
    NRGK_HIP(hipSetDevice(ctx->c.device));
    Chain solver(&ctx->c, mode);
    solver.solve_blocks(d, H, w, U);
  });
}

}  // extern "C"
