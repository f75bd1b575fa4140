// K8: operator rotation into the new eigenbasis.
//
// Wilson-site f-dagger (nrgcore::set_current_fdag_operator, nrgcore/include/nrgcore/nrgcore.hpp:401-477):
//   T(i,j)[kidx + it + il*k, kidx' + it + il'*k] = F_bath[p](bb, bb')(il, il')   (same ii, no chi)
//   f_new(i,j) = U_i^T T U_j.
// Only the kept x kept part is produced (the next assembly reads F_prev(it, it') for kept it, it'
// only, nrgcore.hpp:199-205): Y = T U_j[:, :k_j] is a gather of at most d_bath rows of U_j per row
// (k_fdag_gather), then f_new = U_i[:, :k_i]^T Y is a batched GEMM.
//
// Observables (update_system_operator, nrgcore/include/nrgcore/sysOperator.hpp:35-133):
//   c(i,j)[kidx + it + k_ii*il, kidx' + it' + k_ii'*il] = O_prev(ii, ii')(it, it')  where bb == bb'
//   O_new(i,j) = U_i^T c U_j (full n_i x n_j).
// k_obs_scatter writes the nonzero sub-blocks of c; the two products are batched GEMMs.
#include "common.hpp"

namespace nrgk {


__global__ void __launch_bounds__(256) k_fdag_gather(const RotGroup* __restrict__ groups, const RotSegment* __restrict__ segs,
                                                     const RotTri* __restrict__ tris, const double* __restrict__ U,
                                                     const double* __restrict__ fbath, double* __restrict__ Y) {
  const RotGroup g = groups[blockIdx.x];
  const RotTri tr = tris[g.triple];
  const int rows = g.kii * g.dbb;
  const int kj = tr.kj;
  if (kj == 0) return;
  const double* Uj = U + tr.uj_off;
  for (long long e = threadIdx.x; e < (long long)rows * kj; e += blockDim.x) {
    const int rr = (int)(e / kj), c = (int)(e % kj);
    const int it = rr % g.kii, il = rr / g.kii;
    double v = 0.0;
    for (int s = g.seg_begin; s < g.seg_end; s++) {
      const RotSegment sg = segs[s];
      for (int ilp = 0; ilp < sg.dbbp; ilp++) {
        const double f = fbath[sg.fb_off + il * sg.fb_ld + ilp];
        if (f != 0.0) v += f * Uj[(long long)(sg.kidxp + it + ilp * g.kii) * tr.nj + c];
      }
    }
    Y[tr.y_off + (long long)(g.kidx + rr) * kj + c] = v;
  }
}


__global__ void __launch_bounds__(256) k_obs_scatter(const ObsSeg* __restrict__ segs, const double* __restrict__ Oprev,
                                                     double* __restrict__ C) {
  const ObsSeg s = segs[blockIdx.x];
  const long long tot = (long long)s.kii * s.kiip * s.dbb;
  for (long long e = threadIdx.x; e < tot; e += blockDim.x) {
    const int il = (int)(e / ((long long)s.kii * s.kiip));
    const int rem = (int)(e % ((long long)s.kii * s.kiip));
    const int it = rem / s.kiip, itp = rem % s.kiip;
    C[s.c_off + (long long)(s.kidx + it + s.kii * il) * s.ldc + (s.kidxp + itp + s.kiip * il)] =
        Oprev[s.o_off + (long long)it * s.ldo + itp];
  }
}

void launch_fdag_gather(const RotGroup* g, int ng, const RotSegment* s, const RotTri* t, const double* U,
                        const double* fbath, double* Y, hipStream_t st) {
  if (ng <= 0) return;
  NRGK_LAUNCH(k_fdag_gather, dim3(ng), dim3(256), 0, st, g, s, t, U, fbath, Y);
  NRGK_HIP(hipGetLastError());
}
void launch_obs_scatter(const ObsSeg* s, int ns, const double* Oprev, double* C, hipStream_t st) {
  if (ns <= 0) return;
  NRGK_LAUNCH(k_obs_scatter, dim3(ns), dim3(256), 0, st, s, Oprev, C);
  NRGK_HIP(hipGetLastError());
}

}  // namespace nrgk
