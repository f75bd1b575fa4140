// Launchers of the device kernels (one translation unit per kernel family).
#pragma once
#include "common.hpp"

namespace nrgk {

// k_assemble.hip
void launch_assemble(const AsmSector* sectors, int n_sectors, int n_rows, const AsmSub* subs, const AsmCon* cons,
                     const double* e_prev, const double* e_bath, const double* f_prev, const double* f_bath,
                     double rescale, double* H, hipStream_t s);
void launch_shift_convert(const EigBlock* blocks, const int* row_block, const int* row_local, int n_rows,
                          const double* H, unsigned int* sigma_bits, float* sigma_out, float* W, hipStream_t s);
// k_eig.hip
void launch_jacobi_round(const JacTask* tasks, int n_tasks, float* W, int round, int* flags, double* dgram, int* dgv,
                         float tol, float ang_tol, int max_inner, double jac_far, hipStream_t s);
void launch_normalize_sort(const EigBlock* blocks, int nb, const int4* tiles, int n_tiles, const int4* groups,
                           int n_groups, const float* W, const float* sigma, double* X, double* lam, int* perm,
                           double* nrm_s, double* nrm_g, size_t lds, hipStream_t s);
void launch_oa(const EigBlock* blocks, int nb, const int4* sym_tiles, int n_sym, const int4* all_tiles, int n_all,
               double* S, double* Rp, double* E, double* lam, int* cid, int* sidx, double* offmax, int* stats,
               double* conv, double* scrW, double* scrG, double* scrT, double* rrv, int nmax, double e_tol,
               double d_tol, hipStream_t s);
void launch_copy_converged(const EigBlock* blocks, int nb, const double* conv, double e_tol, double d_tol,
                           double e_pred, double d_pred, const double* src, double* dst, hipStream_t s);
void launch_final_sort(const EigBlock* blocks, int nb, const int4* tiles, int n_tiles, const double* X,
                       const double* lam_in, double* U, double* lam_out, int* perm, size_t lds, hipStream_t s);
size_t normalize_sort_lds(int nmax);
size_t final_sort_lds(int nmax);
void set_kernel_lds_limits(int bytes);
// k_gemm.hip
int prepare_gemm(std::vector<GemmProb>& probs);
void launch_gemm(const GemmProb* d_probs, int n_probs, int tiles, hipStream_t s);
// k_tridiag.hip (the wide blocks)
int ht_grid(const std::vector<HtBlock>& bl, int k, int kind);
void launch_ht_tile_in(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* H, double* A, hipStream_t s);
void launch_ht_first(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* A, double* W, hipStream_t s);
// with events t0, t1 (created, not recorded): the launch records its start and end in them
int launch_ht_symv(const HtBlock* d_bl, const std::vector<HtBlock>& bl, int k, const double* A, const double* V, double* W,
                   hipStream_t s, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
int launch_ht_update(const HtBlock* d_bl, const std::vector<HtBlock>& bl, int k, const double* A, double* V, double* W,
                     hipStream_t s);
int launch_ht_syr2k(const HtBlock* d_bl, const std::vector<HtBlock>& bl, int k, double* A, const double* V, double* W,
                    hipStream_t s);
void launch_ht_stebz(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* W, double* lam, hipStream_t s);
// list: 3 ints per cluster of more than one member (at most the wanted columns), count: one int
void launch_ht_stein(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* W, const double* lam, double* U,
                     double* S, double ctol_rel, int* list, int* count, hipStream_t s);
int ht_stein_slots(int m);  // scratch slots (HtBlock::mpad) of the inverse iteration for m columns
int launch_ht_wy(const HtBlock* d_bl, const std::vector<HtBlock>& bl, int step, const double* V, const double* VT,
                 double* U, double* Ys, hipStream_t s);
void launch_ht_larft(const HtBlock* d_bl, const std::vector<HtBlock>& bl, const double* V, const double* W, double* A,
                     hipStream_t s);
void set_ht_lds_limits(int bytes);
int ht_max_n(int lds_bytes);  // the widest block the tridiagonal kernels hold in LDS
// tune.cpp
std::string tune_cache_path(const Ctx& c);
void tune_load(Ctx& c);
int tune_crossover(Ctx* ctx, const std::string& path, std::vector<std::vector<double>>& table);
// k_trunc.hip
void set_trunc_lds_limit();
void launch_truncate(const int* bn, const int* beoff, int S, const double* lam, long long N_tot, long long K,
                     double errbar, double tie_window, int* kept, TruncOut* out, double* lam_shift, hipStream_t s);
// k_rotate.hip
void launch_fdag_gather(const RotGroup* g, int ng, const RotSegment* s, const RotTri* t, const double* U,
                        const double* fbath, double* Y, hipStream_t st);
void launch_obs_scatter(const ObsSeg* s, int ns, const double* Oprev, double* C, hipStream_t st);
// k_fdm.hip
void launch_fdm_bin(const void* items, int n_items, const double* B, const double* RA, const double* AR,
                    const double* E, const double* bf, double scale, double e_min, double delE, int nbins, double* pos,
                    double* neg, double* spec_sum, double* signed_sum, hipStream_t s);
void launch_partial_trace(const void* items, int n_items, const void* cons, const double* rho, double* red,
                          hipStream_t s);
void launch_thermo(const double* E, const int* shell_off, const double* gs, int n_shells, const double* T, int nT,
                   double* z, double* m1, double* m2, hipStream_t s);
void launch_thermo_ops(const double* E, const double* Q, long long q_stride, int nq, const int* shell_off,
                       const double* gs, int n_shells, const double* T, int nT, double* out, hipStream_t s);
void launch_diag_gather(const DiagItem* items, int n_items, const double* O, double* out, hipStream_t s);
// scratch: broaden_scratch_doubles(n, no) doubles of device memory
size_t broaden_scratch_doubles(int n, int no);
// the allocation-free entry point's fixed device scratch (noalloc.hip)
constexpr size_t NOALLOC_SCRATCH_DOUBLES = size_t(1) << 19;
double* noalloc_scratch_setup(int device);  // with the device current: looks its address up once
double* noalloc_scratch(int device);        // that address (nullptr before the set-up)
void launch_broaden(const double* x, const double* w, int n, const double* omega, int no, double b, double* out,
                    double* scratch, hipStream_t s);
void launch_transpose(const double* in, double* out, int n, hipStream_t s);

}  // namespace nrgk
