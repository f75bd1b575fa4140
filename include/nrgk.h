/* nrgk.h — C ABI of the MI355X NRG iteration engine (libnrgk.so).
 *
 * The reference (srbhp/nrgplusplus) has no C ABI: its boundary is the header-only C++ template
 * API nrgcore<im_type, bath_type> (nrgcore/include/nrgcore/nrgcore.hpp:25-550) plus the free
 * function update_system_operator (nrgcore/include/nrgcore/sysOperator.hpp:35-133) and the
 * observables fdmSpectrum / fdmThermodynamics.  The host C++ layer (nrgplusplus_amd/include/nrgk/)
 * keeps that template API verbatim and calls the entry points below; each entry point names the
 * reference routine it replaces.
 *
 * Conventions
 *  - plain pointers and sizes only; every entry point returns NRGK_OK (0) or a negative status and
 *    never throws; nrgk_last_error() gives the message of the last failure on the calling thread.
 *  - matrices are row-major doubles, as qmatrix<double> (nrgcore/include/utils/qmatrix.hpp:66-75);
 *    eigenvectors are the COLUMNS of a block's U (qmatrix.hpp:390-391).
 *  - one chain is driven by one host thread; chains are independent and may share a context.
 *  - all device work of a chain runs on the chain's own HIP stream.
 */
#ifndef NRGK_H
#define NRGK_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NRGK_OK 0
#define NRGK_ERR_INVALID (-1)   /* bad argument (the reference throws std::runtime_error) */
#define NRGK_ERR_HIP (-2)       /* HIP runtime failure */
#define NRGK_ERR_NOMEM (-3)     /* device capacity exceeded (HBM admission) */
#define NRGK_ERR_STATE (-4)     /* call out of order / empty state */
#define NRGK_ERR_NOCONV (-5)    /* eigensolver did not converge within its sweep budget */
#define NRGK_ERR_UNSUPPORTED (-6)

/* eigensolver / rotation precision modes */
#define NRGK_MODE_MIXED 0   /* FP32 block Jacobi + FP64 Ogita-Aishima refinement, batched  */
#define NRGK_MODE_PLAIN64 1 /* one launch per block: rocSOLVER dsyevd + rocBLAS dgemm, FP64 */

typedef struct nrgk_ctx nrgk_ctx;
typedef struct nrgk_chain nrgk_chain;
typedef struct nrgk_obs nrgk_obs;
typedef struct nrgk_fdm nrgk_fdm;

/* A site model in the reference's duck-typed model concept (nrgcore.hpp:46-48,68-76,119-133):
 * get_basis() -> qn, get_eigenvaluesQ() -> dims/energies, get_chi_Q() -> chi, f_dag_operator -> ops.
 * Operator p has op_nblocks[p] blocks (bi,bj) with data dims[bi] x dims[bj] row-major; blocks of all
 * operators are concatenated in order. */
typedef struct {
  int n_sectors;
  int n_qn;
  const int* qn;          /* [n_sectors * n_qn] */
  const int* dims;        /* [n_sectors] */
  const double* energies; /* [sum dims], ascending within a sector */
  const double* chi;      /* [n_sectors] */
  int n_ops;
  const int* op_nblocks;  /* [n_ops] */
  const int* op_bi;       /* [sum op_nblocks] */
  const int* op_bj;       /* [sum op_nblocks] */
  const double* op_data;  /* concatenated blocks */
} nrgk_site_desc;

/* ---------------------------------------------------------------- site models (host setup) */
typedef struct nrgk_site nrgk_site;
/* spinhalf(eps, U, h) (nrgcore/include/models/spinhalf.hpp:81-124) */
int nrgk_site_spinhalf(double eps, double U, double h, nrgk_site** out);
/* twoChannel() (nrgcore/src/models/twoChannel.cpp:12-48) */
int nrgk_site_two_channel(nrgk_site** out);
/* spinhalfKondo(J, S) (nrgcore/include/models/spinhalfKondo.hpp:28-166) */
int nrgk_site_kondo(double J, double S, nrgk_site** out);
/* f^dagger_sigma chi on a spinhalf site, chi = (-1)^N (examples/fdmSpectrumSiam/main.cpp:86-106);
 * the operators of the returned site are the observable blocks */
int nrgk_site_fdag_chi_spinhalf(nrgk_site** out);
/* the double occupancy n_up n_dn of a spinhalf site (spinhalf::doubleOccupancy / getDoubleOcc,
 * nrgcore/include/models/spinhalf.hpp:106-114, 215): one system operator for the thermodynamic
 * expectation values */
int nrgk_site_double_occupancy_spinhalf(nrgk_site** out);
/* spinless(eps) (nrgcore/include/models/spinnless.hpp): one orbital, charge only, H = eps n */
int nrgk_site_spinless(double eps, nrgk_site** out);
/* spinnlessTwoLead(eps) (nrgcore/include/models/spinlessTwoLead.hpp): two orbitals, charge only,
 * H = eps n_0; both f^dagger couple to the chain (the bath of the resonant-level driver) */
int nrgk_site_spinless_two_lead(double eps, nrgk_site** out);
/* resonantTwoLead(eps, V_L, V_R, U) (nrgcore/include/models/resonantTwoLead.hpp:20-74): a level and
 * the first site of two leads, charge only.  observables = 0: the chain operators (the two lead
 * f^dagger in the eigenbasis); 1: the same sectors with the operators impurityFdag and
 * impurityNparticle, for nrgk_obs_create */
int nrgk_site_resonant_two_lead(double eps, double v_left, double v_right, double U, int observables,
                                nrgk_site** out);
/* spinhalfSC (nrgcore/include/models/spinhalfSC.hpp:27-123): two orbitals, spin only;
 * H = eps (n_up + n_dn) + h (n_up - n_dn) + U n_up n_dn + delta (f_up f_dn + h.c.).  The reference's
 * addSCDelta(delta) replaces H by the pairing term alone: eps = h = U = 0 here. */
int nrgk_site_spinhalf_sc(double eps, double U, double h, double delta, nrgk_site** out);
/* a free two-orbital Wilson site, H = 0, in a symmetry: 0 charge and spin (freeModel.hpp), 1 charge
 * only, 2 spin only (the bath of the superconducting impurities) */
int nrgk_site_free(int symmetry, nrgk_site** out);
/* Any model as the reference's constructors build one (e.g. resonantTwoLead.hpp:31-62): n_states
 * basis states with n_qn quantum numbers each (qn_state[n_states][n_qn]), a symmetric Hamiltonian
 * H[n_states][n_states] that does not couple different quantum numbers, and n_ops operators
 * ops[n_ops][n_states][n_states].  States are grouped into sectors in first-appearance order, each
 * sector diagonalised (eigenvalues ascending), and every nonzero operator block rotated into the
 * eigenbasis, U_i^T F U_j.  chi = (-1)^(sum of the quantum numbers). */
int nrgk_site_from_hamiltonian(int n_states, int n_qn, const int* qn_state, const double* H, int n_ops,
                               const double* ops, nrgk_site** out);
/* The Jordan-Wigner Fock basis of fermionBasis (nrgcore/include/models/fermionBasis.hpp:20-442):
 * f^dagger_i of n_orb orbitals as full 2^n_orb x 2^n_orb matrices (fdag[n_orb][D][D], may be NULL)
 * and the quantum numbers of each basis state for a symmetry (0 charge and spin: 2 numbers;
 * 1 charge only, 2 spin only: 1 number; qn_state[D][n], may be NULL).  n_orb <= 12. */
int nrgk_fock_basis(int n_orb, int symmetry, double* fdag, int* qn_state);
/* impurity.f_dag_operator.push_back(f[0]); push_back(f[1]) (examples/twoChannelSiam/main.cpp:24-25) */
int nrgk_site_duplicate_ops(nrgk_site* site);
/* a view of the site; pointers stay valid until the site is destroyed or modified */
int nrgk_site_view(const nrgk_site* site, nrgk_site_desc* out);
int nrgk_site_destroy(nrgk_site* site);

/* ---------------------------------------------------------------- context */
const char* nrgk_last_error(void);
const char* nrgk_version(void);
/* opens HIP device `device`; no reference counterpart (the reference has no device) */
int nrgk_ctx_create(int device, nrgk_ctx** out);
int nrgk_ctx_destroy(nrgk_ctx* ctx);
/* bytes of device memory the context may hand to chains (0 = 90% of the free HBM) */
int nrgk_ctx_set_capacity(nrgk_ctx* ctx, size_t bytes);
int nrgk_ctx_used_bytes(nrgk_ctx* ctx, size_t* used, size_t* capacity);

/* Autotuner (no reference counterpart): measures, on batches of random symmetric blocks of widths
 * 192..768, the batched Jacobi path (all eigenvectors) against the Householder path (eigenvalues and
 * the first quarter of the eigenvectors) and sets the context's crossover to the widest width at which
 * the Jacobi path is not slower.  The result and the measured times go to a cache file (cache_path, or
 * NULL: $NRGK_TUNE_CACHE, else $HOME/.cache/nrgk/tune-<arch>-<CUs>cu.txt) that every context created
 * later on the same GPU model reads.  table[n_rows][4]: width, blocks, Jacobi ms, Householder ms (may
 * be NULL).  NRGK_MIXED_MAX, when set, overrides the crossover per call. */
int nrgk_tune(nrgk_ctx* ctx, const char* cache_path, int* crossover, double* table, int cap_rows, int* n_rows);
/* the context's crossover and whether it came from the autotuner (its cache or nrgk_tune) */
int nrgk_ctx_crossover(nrgk_ctx* ctx, int* crossover, int* tuned);

/* ---------------------------------------------------------------- chain (nrgcore) */
/* nrgcore(im_type&, bath_type&) (nrgcore.hpp:44-53) + set_parameters(max_kept) (nrgcore.hpp:336).
 * The model-consistency checks of nrgcore::test (nrgcore.hpp:117-134) return NRGK_ERR_INVALID. */
int nrgk_chain_create(nrgk_ctx* ctx, const nrgk_site_desc* impurity, const nrgk_site_desc* bath,
                      size_t max_kept, int mode, nrgk_chain** out);
int nrgk_chain_destroy(nrgk_chain* chain);
/* nrgcore::set_parameters (nrgcore.hpp:336-342): resets the iteration counters */
int nrgk_chain_set_parameters(nrgk_chain* chain, size_t max_kept);
/* replaces the bath model read on every add_bath_site (nrgcore.hpp:68-70) */
int nrgk_chain_set_bath(nrgk_chain* chain, const nrgk_site_desc* bath);
/* nrgcore::add_bath_site (nrgcore.hpp:65-97): basis, sector assembly, batched eigensolve.
 * The Wilson-site f-dagger rotation (nrgcore.hpp:91, set_current_fdag_operator 401-477) is
 * deferred to update_internal_state and restricted to kept x kept blocks, the only part the next
 * assembly reads (nrgcore.hpp:199-205). */
int nrgk_chain_add_bath_site(nrgk_chain* chain, const double* t_hopping, int n_hopping, double rescale);
/* nrgcore::update_internal_state (nrgcore.hpp:105-109): truncation (discard_higher_energies,
 * nrgcore.hpp:269-327), energy shift, f-dagger rotation */
int nrgk_chain_update_internal_state(nrgk_chain* chain);
/* nrgk_chain_add_bath_site + nrgk_chain_update_internal_state in one call */
int nrgk_chain_step(nrgk_chain* chain, const double* t_hopping, int n_hopping, double rescale);
/* qmatrix::diag (nrgcore/include/utils/qmatrix.hpp:375-404: LAPACKE_dsyevd(ROW_MAJOR, 'V', 'U', n,
 * A, n, w), A becomes the eigenvectors) for n_blocks symmetric blocks in one call, through the same
 * batched eigensolver as a chain's sectors.  Block b is dims[b] x dims[b], row-major; the blocks are
 * packed one after another in H (sum dims[b]^2 doubles, only the blocks' full storage is read).
 * Ascending eigenvalues go to w (sum dims[b]), the eigenvectors as the columns of each block to U
 * (packed like H).  Host arrays; mode NRGK_MODE_MIXED or NRGK_MODE_PLAIN64.  Allocates its
 * device buffers per call.  NRGK_ERR_NOCONV when a block does not converge (the FP32 Jacobi, the
 * FP64 refinement within its step limit, or dsyevd); the reference only prints dsyevd's info. */
int nrgk_eigh_batched(nrgk_ctx* ctx, int n_blocks, const int* dims, const double* H, double* w, double* U,
                      int mode);
/* Several independent chains (z-shifted discretisations, parameter points, requests) advanced by one
 * Wilson site each, concurrently: one host worker and the chain's own stream per chain, so that the
 * chains' launches overlap on the GPU.  No counterpart in the reference, whose drivers step one
 * chain; chain i runs that driver loop (examples/fdmSpectrumSiam/main.cpp:86-95):
 *   add_bath_site(t + i*n_hopping, n_hopping, rescale[i]);
 *   update_system_operator on obs[i*n_obs .. i*n_obs + n_obs - 1];
 *   if checkHigherEnergyDiscarded() and recorders (may be NULL) has recorders[i] != NULL:
 *     record the shell into recorders[i] (rawData.saveCurrentData), recorded[i] = 1, else 0;
 *   update_internal_state().
 * Chains must be distinct, each observable and recorder must belong to its chain.  n_threads <= 0:
 * min(n_chains, 16) workers.  recorded may be NULL.  Every chain completes its step; the first
 * failing chain's status is returned, with its index in nrgk_last_error(). */
int nrgk_chains_step(nrgk_chain* const* chains, int n_chains, const double* t_hopping, int n_hopping,
                     const double* rescale, nrgk_obs* const* obs, int n_obs, nrgk_fdm* const* recorders,
                     int n_threads, int* recorded);
/* nrg_iterations_cnt / nrg_iterations_min (nrgcore.hpp:536-538) */
int nrgk_chain_counters(nrgk_chain* chain, int* cnt, int* cnt_min);
/* overwrite nrg_iterations_cnt (drivers decrement it, examples/fdmSpectrumSiam/main.cpp:163) */
int nrgk_chain_set_counter(nrgk_chain* chain, int cnt);
/* checkHigherEnergyDiscarded (nrgcore.hpp:384-386) */
int nrgk_chain_higher_energy_discarded(nrgk_chain* chain, int* out);
/* all_eigenvalue (nrgcore.hpp:520): sorted, pre-shift.  *n receives the length; out may be NULL */
int nrgk_chain_all_eigenvalues(nrgk_chain* chain, double* out, size_t cap, size_t* n);
/* relativeGroundStateEnergy (nrgcore.hpp:522) */
int nrgk_chain_relative_gs(nrgk_chain* chain, double* out, size_t cap, size_t* n);
/* current_sysmQ (nrgcore.hpp:528) / eigenvaluesQ sizes: qn[n_sectors*n_qn], dims[n_sectors];
 * both may be NULL to query n_sectors */
int nrgk_chain_sectors(nrgk_chain* chain, int* n_sectors, int* qn, int* dims);
/* pre_sysmQ (nrgcore.hpp:530): the sectors of the previous (kept) basis, qn[n_sectors*n_qn] (may be NULL);
 * after update_internal_state they are the current ones, as in the reference (nrgcore.hpp:105-109) */
int nrgk_chain_pre_sectors(nrgk_chain* chain, int* n_sectors, int* qn);
/* coupled_nQ_index (nrgcore.hpp:534, create_next_basis 238-261): for each current sector, the list of
 * i * n_bath_sectors + j over (previous sector i, bath sector j), first-seen order.  counts[n_sectors]
 * (may be NULL), flat[cap] (may be NULL), total: the summed length */
int nrgk_chain_coupled_index(nrgk_chain* chain, int* counts, int* flat, size_t cap, size_t* total);
/* eigenvaluesQ (nrgcore.hpp:532), concatenated over sectors (shifted after truncation) */
int nrgk_chain_eigenvalues_q(nrgk_chain* chain, double* out, size_t cap, size_t* n);
/* eigenvaluesQ_kept_indices sizes (nrgcore.hpp:524); kept indices are always a prefix */
int nrgk_chain_kept_counts(nrgk_chain* chain, int* out, int cap, int* n_sectors);
/* current_hamiltonQ[sector] (nrgcore.hpp:526): U after the eigensolve, n x n row-major */
int nrgk_chain_block_u(nrgk_chain* chain, int sector, double* out, size_t cap);
/* the assembled sector Hamiltonian before the eigensolve (diagnostic; needs keep_h) */
int nrgk_chain_keep_hamiltonians(nrgk_chain* chain, int on);
int nrgk_chain_block_h(nrgk_chain* chain, int sector, double* out, size_t cap);
/* getPreWilsonSiteOperators (nrgcore.hpp:548): kept x kept blocks of f-dagger op p after
 * update_internal_state.  bi/bj/rows/cols receive the block table; data the concatenated blocks. */
int nrgk_chain_fdag_blocks(nrgk_chain* chain, int op, int* n_blocks, int* bi, int* bj, int* rows,
                           int* cols, double* data, size_t cap);
/* Overwrite the state the next add_bath_site reads (what NrgData::loadCurrentData does for the
 * members it restores, nrgcore/include/nrgcore/nrgData.hpp:266-293): previous sectors (qn),
 * their kept energies and the kept x kept f-dagger blocks.  Layout as nrgk_site_desc with
 * dims = kept counts. */
int nrgk_chain_set_state(nrgk_chain* chain, const nrgk_site_desc* state, int cnt);
/* per-step device timings (ms): assemble, eigensolve, truncate, rotate; and eigensolver stats */
int nrgk_chain_step_stats(nrgk_chain* chain, double* ms4, int* jacobi_sweeps, int* max_cluster);
/* the last truncation: E_cut = sorted[K] (discard_higher_energies, nrgcore.hpp:285) and the number of
 * eigenvalues within 2e-10 of E_cut + 1e-8 (either side), whose keep/discard decision an eigenvalue
 * error of that size could flip (0 when nothing was truncated); no reference counterpart */
int nrgk_chain_truncation_stats(nrgk_chain* chain, double* e_cut, int* near_ties);
/* kernel launches (a rocBLAS/rocSOLVER call counts once) and host stream synchronisations of the
 * last step, from add_bath_site to the end of update_internal_state (observable updates and shell
 * records made between them included); no reference counterpart */
int nrgk_chain_step_counts(nrgk_chain* chain, long long* launches, long long* syncs);
/* refinement statistics of the last eigensolve: Ogita-Aishima steps taken and blocks that did not
 * meet the convergence test within the step budget */
int nrgk_chain_refine_stats(nrgk_chain* chain, int* oa_steps, int* unconverged);
/* Per-kernel-family device timing, the counterpart of the reference's RAII wall-clock timer
 * (nrgcore/include/utils/timer.hpp:22-79) at kernel granularity.  While enabled (enabling clears the
 * totals), each run of consecutive launches of one family is bracketed by HIP events on the chain's
 * stream, and its device time is accumulated with the launch count and the algorithmic flops and
 * bytes of those launches.  Families: "jacobi_round", "oa_update", "gemm_f64". */
typedef struct {
  char family[32];
  double ms;
  long long launches;
  double flops;
  double bytes;
} nrgk_kernel_stat;
int nrgk_chain_profile(nrgk_chain* chain, int enable);
/* restricts the timing to one family (NULL or "": all), so that a caller timing one kernel records
 * one event pair per run of that family only */
int nrgk_chain_profile_only(nrgk_chain* chain, const char* family);
int nrgk_chain_kernel_stats(nrgk_chain* chain, nrgk_kernel_stat* out, int cap, int* n_out);

/* ---------------------------------------------------------------- observables (sysOperator.hpp) */
/* A set of system operators (e.g. d^dagger chi) defined on the impurity sectors; blocks given as
 * in nrgk_site_desc (op_nblocks/op_bi/op_bj/op_data, dims from the impurity). */
int nrgk_obs_create(nrgk_chain* chain, int n_ops, const int* op_nblocks, const int* op_bi,
                    const int* op_bj, const double* op_data, nrgk_obs** out);
int nrgk_obs_destroy(nrgk_obs* obs);
/* update_system_operator (sysOperator.hpp:35-133): call after add_bath_site, before
 * update_internal_state.  Produces full n_i x n_j blocks in the new eigenbasis. */
int nrgk_obs_update(nrgk_obs* obs);
int nrgk_obs_blocks(nrgk_obs* obs, int op, int* n_blocks, int* bi, int* bj, int* rows, int* cols,
                    double* data, size_t cap);

/* ---------------------------------------------------------------- FDM spectrum (fdmSpectrum.hpp) */
/* Recorder of the forward pass (the state NrgData saves per iteration, nrgData.hpp:134-205), kept
 * resident in HBM.  nrgk_fdm_record is called where the reference drivers call saveCurrentData +
 * saveqOperator (examples/fdmSpectrumSiam/main.cpp:134-137): after add_bath_site and
 * nrgk_obs_update, before update_internal_state. */
int nrgk_fdm_create(nrgk_chain* chain, nrgk_obs* obs, int n_bins, double e_min, double e_max,
                    nrgk_fdm** out);
int nrgk_fdm_destroy(nrgk_fdm* fdm);
int nrgk_fdm_record(nrgk_fdm* fdm);
/* The T = 0 backward pass (fdmSpectrum::calcSpectrum, fdmSpectrum.hpp:55-63, driven as
 * examples/fdmSpectrumSiam/main.cpp:158-180) over the recorded shells with iteration counter
 * cnt >= first_cnt, last shell first.  A shell recorded at counter cnt uses the energy scale
 * lambda^{-(cnt + scale_offset)/2} (the SIAM drivers: scale_offset = -1).  pos/neg: [n_ops][n_bins],
 * weights summed over shells; traces/sums: per processed shell (rhoTrace, specSum), may be NULL. */
int nrgk_fdm_backward(nrgk_fdm* fdm, double lambda, int scale_offset, int first_cnt, double* pos,
                      double* neg, double* traces, double* sums, int cap_shells, int* n_shells);
/* NrgData::saveCurrentData + saveqOperator (nrgcore/include/nrgcore/nrgData.hpp:134-205) of every
 * recorded shell, in an HDF5 file of the reference's layout (h5stream.hpp:244-478): /NrgItr{cnt}/
 * current_hamiltonQ{s}, current_sysmQ{s}, pre_sysmQ{s}, eigenvaluesQ{s}, coupled_nQ_index{s},
 * eigenvaluesQ_kept_indices{s}; /Eigenvalues{cnt}; and, when the recorder has an observable and
 * op_group is not NULL, {op_group}{cnt-1} with attribute operatorSize and at{p}/qmat{k},
 * idVector{k}, colVector, rowVector.  The HDF5 library is loaded when first needed. */
int nrgk_fdm_save(nrgk_fdm* rec, const char* path, const char* op_group);
/* h5stream::h5stream(name) (nrgcore/include/utils/h5stream.hpp:244-478): creates (truncates) an HDF5
 * file; nrgk_h5_write adds a one-dimensional dataset of n doubles to it (its groups created as
 * needed), as h5stream::write(std::vector<double>, name) does.  The HDF5 library is loaded when first
 * needed. */
int nrgk_h5_create(const char* path);
int nrgk_h5_write(const char* path, const char* name, const double* data, size_t n);
/* NrgData::loadCurrentData + loadqOperator (nrgcore/include/nrgcore/nrgData.hpp:214-293) for every
 * shell of a file nrgk_fdm_save wrote: the recorder's shells are replaced by the file's, rebuilt on the
 * device (U blocks, eigenvalues, sectors, coupled index, kept counts and, when the recorder has an
 * observable, its blocks from {op_group}{cnt-1}), so that the backward pass or the thermodynamics runs
 * in a later process than the forward pass.  The recorder's chain must have the file's bath model. */
int nrgk_fdm_load(nrgk_fdm* rec, const char* path, const char* op_group);
/* a one-dimensional dataset of an HDF5 file as doubles (n: its length; out may be NULL) */
int nrgk_h5_read(const char* path, const char* dataset, double* out, size_t cap, size_t* n);
/* Per operator, the signed weights w+ + w- of the last nrgk_fdm_backward summed over its shells,
 * taken before the |.|, the 1e-20 cut and the bin-range test of rhoDotOperators
 * (fdmSpectrum.hpp:231-263).  Over a complete basis of discarded states this is <{B, B+}> (1 for
 * d+ chi): the FDM sum rule, which the reference does not compute.  n_ops receives the operator
 * count; out[cap] may be NULL. */
int nrgk_fdm_signed_weights(nrgk_fdm* fdm, double* out, int cap, int* n_ops);
/* The reference's backward loop one shell at a time (examples/fdmSpectrumSiam/main.cpp:158-180,
 * examples/entropySIAM/main.cpp:80-91): after the driver decrements nrg_iterations_cnt and
 * NrgData::loadCurrentData restores the shell, fdmSpectrum::calcSpectrum(scale) or
 * fdmThermodynamics::calcThermodynamics(scale) processes it.  nrgk_fdm_add_shell registers the shell
 * recorded at iteration counter cnt with its energy scale; the registered shells are processed in
 * registration order (the last shell of the chain first) by nrgk_fdm_run_spectrum (the FDM pass,
 * outputs as nrgk_fdm_backward) or nrgk_thermo_run (fdmThermodynamics::saveFinalData: entropy S,
 * specific heat C and, when a system operator was set, its expectation values ops[n_ops][nT], may be
 * NULL), which then clear the registration.  nrg_max: nrg_iterations_cnt when the reference's
 * fdmThermodynamics object was constructed (nrgMaxIterations, fdmThermodynamics.hpp:26-29). */
int nrgk_fdm_add_shell(nrgk_fdm* fdm, int cnt, double scale);
int nrgk_fdm_run_spectrum(nrgk_fdm* fdm, double* pos, double* neg, double* traces, double* sums, int cap_shells,
                          int* n_shells);
int nrgk_thermo_run(nrgk_fdm* fdm, int nrg_max, const double* T, int nT, double* S, double* C, double* ops);
/* bin centres e_min*exp(i/delE) (fdmSpectrum.hpp:323-327) */
int nrgk_fdm_energy_points(nrgk_fdm* fdm, double* out);

/* ---------------------------------------------------------------- thermodynamics */
/* fdmThermodynamics::calcThermodynamics + saveFinalData (static/fdmThermodynamics.hpp:42-236) over
 * the shells recorded by an nrgk_fdm recorder (created with obs = NULL for a thermodynamics-only
 * record), shells with cnt >= first_cnt, energy scale lambda^{-(cnt + scale_offset)/2}
 * (examples/entropySIAM/main.cpp:83-84: scale_offset = -1; examples/entropyKondo/main.cpp:105:
 * scale_offset = 0).  Writes entropy S[nT] and specific heat C[nT]. */
int nrgk_thermo(nrgk_fdm* rec, double lambda, int scale_offset, int first_cnt, const double* T,
                int nT, double* S, double* C);
/* fdmThermodynamics::setSystemOperator (static/fdmThermodynamics.hpp:39-41): from the next
 * nrgk_fdm_record on, each recorded shell also keeps the diagonal of the (i, i) blocks of the
 * operators of sysop in that shell's eigenbasis (calcThermodynamics, fdmThermodynamics.hpp:63-87).
 * sysop must belong to the recorder's chain and be updated (nrgk_obs_update) before each record. */
int nrgk_fdm_set_system_operator(nrgk_fdm* rec, nrgk_obs* sysop);
/* qsysOPeratorExpValue of saveFinalData (fdmThermodynamics.hpp:191-211): the thermal expectation
 * value of each system operator, sum_m w_m sum_{l discarded in m} exp((gs_m - E_l)/T)/Z'_m O_ll,
 * with the shell weights and energies of nrgk_thermo.  out: [n_ops][nT].  Every shell used must
 * have been recorded after nrgk_fdm_set_system_operator. */
int nrgk_thermo_expectation(nrgk_fdm* rec, double lambda, int scale_offset, int first_cnt, const double* T,
                            int nT, double* out);

/* ---------------------------------------------------------------- z-shifted discretisation */
/* The Wilson chain of a bath discretised on the z-shifted logarithmic mesh (Campo-Oliveira levels),
 * the reference's z-averaging generator: calcTn (independFiles/zAverage/hoppingElement.cpp:16-105,
 * flat band) and zAverage (independFiles/zAverage/zAverage.hpp:12-195, tabulated density of states).
 * Host computation in quadruple precision with full reorthogonalisation (the reference: 1000-digit
 * MPFR, plain three-term Lanczos).
 *   z: twist in (0, 1] (z = 0 is the same mesh as z = 1); lambda > 1; D: half bandwidth
 *   n_x = 0: flat band rho = 1/(2D); else rho tabulated at ascending x[0..n_x), linear between points
 *   quad: 0 exact integrals of that interpolant; 1 the reference's 100-point left sums
 *         (zAverage/quadrature.hpp:13-26)
 *   n_intervals: intervals per half band (0: 2 n_sites + 20; calcTn uses 11)
 *   normalize: 1 normalise the coupling vector (zAverage.hpp:131-143); 0 keep it (calcTn)
 * Outputs t_n Lambda^{n/2} and eps_n Lambda^{n/2}, n < n_sites (getHoppingElement's scaling,
 * zAverage.hpp:33-43); entries past the star's m - 1 hoppings are 0.  No GPU work. */
int nrgk_zdisc_chain(double z, double lambda, double D, const double* x, const double* rho, int n_x, int quad,
                     int n_intervals, int normalize, int n_sites, double* t_out, double* eps_out);

/* ---------------------------------------------------------------- request loop (serving) */
/* Impurity solves served concurrently: no reference counterpart (the reference's drivers run one
 * chain per process).  A request is the FDM spectral-function driver of
 * examples/fdmSpectrumSiam/main.cpp:86-180 for its own models and parameters: the first site with
 * hopping first_hopping (rescale 1), then n_sites Wilson sites hopping(n, lambda) (rescale sqrt(lambda)
 * from the second), the observable's operators rotated each step, a shell recorded whenever higher
 * energies were discarded, the T = 0 FDM backward pass on n_bins bins over [1e-10, 10], and the weights
 * summed over the operators and broadened (log-Gaussian, width broadening; spec.py:27-44) on the n_omega
 * positive frequencies omega: A_pos(omega) from the positive-frequency weights, A_neg(omega) (i.e.
 * A(-omega)) from the negative ones.  fail_at_site >= 0 makes that site's hopping NaN (failure
 * injection: the request fails with NRGK_ERR_INVALID, the others are not affected); -1 otherwise.
 * The model descriptions and omega are copied at submission. */
typedef struct {
  const nrgk_site_desc* impurity;
  const nrgk_site_desc* bath;
  const nrgk_site_desc* observable; /* operators on the impurity sectors (op_nblocks/op_bi/op_bj/op_data) */
  size_t max_kept;
  double lambda;
  double first_hopping;
  int n_sites;
  int n_bins;
  const double* omega;
  int n_omega;
  double broadening;
  int fail_at_site;
} nrgk_request;
typedef struct nrgk_server nrgk_server;
/* A server on a context with n_slots requests running at once, each admitted only when its estimated
 * device footprint (chain arenas + one FDM shell per site) fits beside the running ones within
 * hbm_budget bytes (0: no budget of its own; the context's capacity still fails allocations with
 * NRGK_ERR_NOMEM).  Admission is first come, first served; every running request has its own chain and
 * stream, so their iterations interleave on the device, and a finished request frees its slot and
 * memory for the next at once. */
int nrgk_server_create(nrgk_ctx* ctx, int n_slots, size_t hbm_budget, nrgk_server** out);
/* waits for the requests already submitted, then stops the workers */
int nrgk_server_destroy(nrgk_server* srv);
/* queues a request (returns at once); NRGK_ERR_NOMEM if its footprint exceeds the budget */
int nrgk_server_submit(nrgk_server* srv, const nrgk_request* req, long long* id);
/* *done = 1 once request id has finished (successfully or not) */
int nrgk_server_poll(nrgk_server* srv, long long id, int* done);
/* Waits for request id and returns its status (NRGK_OK or the failure of its solve, message in
 * nrgk_last_error()); A_pos and A_neg receive n_omega values each (may be NULL); info (may be NULL)
 * receives 6 doubles: latency ms (submission to completion), queue ms (submission to admission),
 * total FDM weight, iterations, recorded shells, admitted footprint in bytes.  The request is released. */
int nrgk_server_wait(nrgk_server* srv, long long id, double* A_pos, double* A_neg, double* info);
/* running and queued requests, bytes reserved by the running ones, the most that ran at once and the
 * most bytes they reserved at once (every pointer may be NULL) */
int nrgk_server_stats(nrgk_server* srv, int* running, int* queued, size_t* reserved, int* peak_running,
                      size_t* peak_reserved);

/* ---------------------------------------------------------------- broadening */
/* log-Gaussian broadening (examples/fdmSpectrumSiam/spec.py:27-44): out[o] = (1/w_o) sum_k
 * weight_k exp(-(ln(w_o/x_k)/b - b/4)^2)/(b sqrt(pi)), on the device */
int nrgk_broaden(nrgk_ctx* ctx, const double* x, const double* weight, int n, const double* omega,
                 int n_omega, double b, double* out);
/* No counterpart in the reference: nrgk_broaden for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  x, weight, omega and out are DEVICE buffers of the
 * caller; the kernels are enqueued on `stream` (a hipStream_t, NULL for the default stream) and the
 * call allocates nothing: the partial sums use a fixed 4 MiB device array per GPU, so
 * ceil(n / 512) * n_omega must not exceed 524288 (else NRGK_ERR_INVALID).  Results equal
 * nrgk_broaden's bit for bit.  Calls that overlap on one GPU race on that array. */
int nrgk_broaden_noalloc(nrgk_ctx* ctx, const double* x, const double* weight, int n, const double* omega,
                         int n_omega, double b, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* NRGK_H */
