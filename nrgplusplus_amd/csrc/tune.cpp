// Autotuning of the eigensolver's crossover (no counterpart in the reference): the block width above
// which a sector block leaves the batched FP32 Jacobi (+ FP64 refinement) for the FP64 Householder
// path.  For each candidate width n, a batch of random symmetric n x n blocks is solved both ways on
// the device -- the Jacobi path forming every eigenvector (the refinement needs them), the Householder
// path the eigenvalues and the first n/4 eigenvectors (the kept columns an RG step reads) -- and the
// crossover is the widest candidate at which the Jacobi path is not slower.  The result is cached in a
// text file per GPU model and CU count, which every context reads when it is created.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "chain.hpp"
#include "kernels.hpp"

namespace nrgk {

std::string tune_cache_path(const Ctx& c) {
  if (const char* p = std::getenv("NRGK_TUNE_CACHE")) return p;
  const char* home = std::getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/nrgk/tune-" + c.arch + "-" + std::to_string(c.n_cu) + "cu.txt";
}

// "crossover <n>" from the cache file, if it exists and was written for this GPU
void tune_load(Ctx& c) {
  std::ifstream f(tune_cache_path(c));
  if (!f) return;
  std::string line, key;
  std::string arch;
  int cu = -1, cross = -1;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    is >> key;
    if (key == "gpu") is >> arch >> cu;
    else if (key == "crossover") is >> cross;
  }
  if (arch == c.arch && cu == c.n_cu && cross >= 32 && cross <= 768) {
    c.mixed_max = cross;
    c.tuned = true;
  }
}

// one timed solve of the blocks: eigenvalues of all, eigenvectors of all (Jacobi) or of the first
// quarter of the columns (Householder), ms of device work between two stream synchronisations
static double timed_solve(Ctx* ctx, const std::vector<int>& dims, const std::vector<double>& H, int mixed_max) {
  Chain c(ctx, NRGK_MODE_MIXED);
  c.mixed_max_override = mixed_max;
  std::vector<double> w, U;
  double best = 1e300;
  for (int rep = 0; rep < 2; rep++) {
    c.prepare_blocks(dims, H.data());
    NRGK_HIP(stream_sync(c.st));
    const auto t0 = std::chrono::steady_clock::now();
    c.eigensolve();
    if (!c.ht.empty()) {
      c.kept.assign(dims.size(), 0);
      for (size_t s = 0; s < dims.size(); s++) c.kept[s] = dims[s] / 4;
      c.wide_vectors(false);
    }
    c.sync();
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    best = std::min(best, ms);
  }
  return best;
}

int tune_crossover(Ctx* ctx, const std::string& path, std::vector<std::vector<double>>& table) {
  const int cand[] = {192, 256, 320, 384, 448, 512, 576, 640, 704, 768};
  unsigned long long st = 0x2545F4914F6CDD1Dull;
  auto rnd = [&] {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  };
  int cross = cand[0];
  table.clear();
  for (int n : cand) {
    const int nb = std::max(4, 16384 / n);
    std::vector<int> dims(nb, n);
    std::vector<double> H((size_t)nb * n * n);
    for (int b = 0; b < nb; b++) {
      double* A = H.data() + (size_t)b * n * n;
      for (int i = 0; i < n; i++)
        for (int j = 0; j <= i; j++) {
          const double v = i == j ? 4.0 * rnd() + 0.01 * i : rnd() / std::sqrt((double)n);
          A[(size_t)i * n + j] = A[(size_t)j * n + i] = v;
        }
    }
    const double tj = timed_solve(ctx, dims, H, n);
    const double th = timed_solve(ctx, dims, H, n - 1);
    table.push_back({(double)n, (double)nb, tj, th});
    if (tj <= th) cross = n;
  }
  ctx->mixed_max = cross;
  ctx->tuned = true;
  if (!path.empty()) {
    const size_t sl = path.find_last_of('/');
    if (sl != std::string::npos) {
      const std::string dir = path.substr(0, sl);
      std::string cmd = "mkdir -p '" + dir + "'";
      (void)std::system(cmd.c_str());
    }
    std::ofstream f(path);
    if (f) {
      f << "gpu " << ctx->arch << " " << ctx->n_cu << "\n";
      f << "crossover " << cross << "\n";
      for (auto& r : table)
        f << "# n " << (int)r[0] << " blocks " << (int)r[1] << " jacobi_ms " << r[2] << " householder_ms " << r[3] << "\n";
    }
  }
  return cross;
}

}  // namespace nrgk
