// Oracle-only: run-time binding of the host LAPACK/CBLAS (dsyevd, dgemm).
//
// This is test infrastructure.  The reference calls LAPACKE_dsyevd(ROW_MAJOR,'V','U') for every
// eigensolve (nrgcore/include/utils/qmatrix.hpp:390-391) and cblas_dgemm(RowMajor,NoTrans,NoTrans)
// for every product (qmatrix.hpp:364-365).  The image ships no LAPACKE/CBLAS headers, only the
// libraries, so the two entry points are bound with dlopen/dlsym.  The search order is MKL's
// single dynamic library, then the OpenBLAS bundled in the scipy wheel.  When neither loads the
// oracle throws: it never falls back to a home-made solver.
#pragma once
#include <dlfcn.h>

#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

namespace nrg_oracle {

struct Lapack {
  // LAPACKE / CBLAS constants (values fixed by the published LAPACKE and CBLAS interfaces).
  static constexpr int kRowMajor = 101;
  static constexpr int kNoTrans = 111;
  static constexpr int kTrans = 112;

  using dsyevd_t = int (*)(int, char, char, int, double*, int, double*);
  using dgemm_t = void (*)(int, int, int, int, int, int, double, const double*, int,
                           const double*, int, double, double*, int);
  using set_threads_t = void (*)(int);

  dsyevd_t dsyevd = nullptr;
  dgemm_t dgemm = nullptr;
  set_threads_t set_threads = nullptr;
  std::string library;

  static Lapack& get() {
    static Lapack inst;
    return inst;
  }

 private:
  Lapack() {
    std::vector<std::string> cands;
    if (const char* env = std::getenv("NRG_ORACLE_LAPACK")) cands.emplace_back(env);
    cands.emplace_back("/opt/conda/lib/libmkl_rt.so.1");
    cands.emplace_back("/opt/conda/lib/libmkl_rt.so");
    cands.emplace_back("libmkl_rt.so.1");
    cands.emplace_back("libmkl_rt.so");
    for (const auto& c : cands) {
      void* h = dlopen(c.c_str(), RTLD_NOW | RTLD_GLOBAL);
      if (!h) continue;
      dsyevd = reinterpret_cast<dsyevd_t>(dlsym(h, "LAPACKE_dsyevd"));
      dgemm = reinterpret_cast<dgemm_t>(dlsym(h, "cblas_dgemm"));
      set_threads = reinterpret_cast<set_threads_t>(dlsym(h, "MKL_Set_Num_Threads"));
      if (dsyevd && dgemm) {
        library = c;
        return;
      }
    }
    // scipy-bundled OpenBLAS exports the same routines under a scipy_ prefix.
    const char* globs[] = {"/usr/local/lib/python3.10/dist-packages/scipy.libs/",
                           "/usr/lib/python3/dist-packages/scipy.libs/"};
    for (const char* dir : globs) {
      std::string cmd = std::string(dir);
      void* h = nullptr;
      // the file name carries a build hash; try the one shipped in this image
      for (const char* f : {"libscipy_openblas-68440149.so"}) {
        h = dlopen((cmd + f).c_str(), RTLD_NOW | RTLD_GLOBAL);
        if (h) {
          dsyevd = reinterpret_cast<dsyevd_t>(dlsym(h, "scipy_LAPACKE_dsyevd"));
          dgemm = reinterpret_cast<dgemm_t>(dlsym(h, "scipy_cblas_dgemm"));
          set_threads = reinterpret_cast<set_threads_t>(dlsym(h, "scipy_openblas_set_num_threads"));
          if (dsyevd && dgemm) {
            library = cmd + f;
            return;
          }
        }
      }
    }
    throw std::runtime_error(
        "nrg_oracle: no host LAPACK found (tried MKL libmkl_rt and scipy OpenBLAS); "
        "set NRG_ORACLE_LAPACK to a library exporting LAPACKE_dsyevd and cblas_dgemm");
  }
};

}  // namespace nrg_oracle
