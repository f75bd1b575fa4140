// The RG flow of examples/rgflowSIAM/main.cpp, written against the nrgcore C++ API of
// include/nrgk/nrgcore.hpp: spinhalf impurity, spinhalf bath, first site {V, V} at rescale 1, then
// {t_n, t_n} at rescale 1 for n = 0 and sqrt(Lambda) after.  Prints one line per iteration: the iteration number, the
// state count and all_eigenvalue (first n_print values and the largest).  With n_copies > 1 that many
// identical chains are stepped together (nrgk::step_chains, concurrently); the first is printed and
// the others must agree with it to the last bit.
//
// usage: nrgk_rgflow_siam eps U Gamma Lambda K n_sites [n_print [n_copies]]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "../../include/nrgk/nrgcore.hpp"

int main(int argc, char** argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s eps U Gamma Lambda K n_sites [n_print [n_copies]]\n", argv[0]);
    return 2;
  }
  const double eps = std::atof(argv[1]), U = std::atof(argv[2]), Gamma = std::atof(argv[3]);
  const double Lambda = std::atof(argv[4]);
  const size_t K = (size_t)std::atol(argv[5]);
  const int n_sites = std::atoi(argv[6]);
  const size_t n_print = argc > 7 ? (size_t)std::atol(argv[7]) : 8;
  const int n_copies = argc > 8 ? std::atoi(argv[8]) : 1;
  if (n_copies < 1) {
    std::fprintf(stderr, "n_copies must be at least 1\n");
    return 2;
  }
  try {
    nrgk::site impurity = nrgk::site::spinhalf(eps, U);
    nrgk::site bath = nrgk::site::spinhalf(0.0, 0.0);
    std::vector<std::unique_ptr<nrgk::nrgcore<>>> chains;
    std::vector<nrgk::nrgcore<>*> all;
    for (int c = 0; c < n_copies; c++) {
      chains.push_back(std::make_unique<nrgk::nrgcore<>>(impurity, bath));
      chains.back()->set_parameters(K);
      all.push_back(chains.back().get());
    }
    nrgk::nrgcore<>& chain = *chains[0];
    auto step = [&](double t, double rescale) {
      if (n_copies == 1) {
        chain.add_bath_site({t, t}, rescale);
        chain.update_internal_state();
      } else {
        nrgk::step_chains(all, std::vector<std::vector<double>>(n_copies, {t, t}),
                          std::vector<double>(n_copies, rescale));
      }
    };
    const double V = std::sqrt(2.0 * Gamma / M_PI);
    step(V, 1.0);
    for (int n = 0; n <= n_sites; n++) {
      if (n > 0) {
        const double t = nrgk::hopping((size_t)(n - 1), Lambda);
        step(t, n > 1 ? std::sqrt(Lambda) : 1.0);  // examples/rgflowSIAM/main.cpp:33-42
      }
      const auto& e = chain.all_eigenvalue;
      for (int c = 1; c < n_copies; c++)
        if (chains[c]->all_eigenvalue != e) {
          std::fprintf(stderr, "chain %d differs from chain 0 at iteration %d\n", c, n);
          return 3;
        }
      std::printf("%d %zu", n, e.size());
      for (size_t i = 0; i < e.size() && i < n_print; i++) std::printf(" %.17g", e[i]);
      std::printf(" %.17g\n", e.empty() ? 0.0 : e.back());
    }
    // qmatrix::diag through the batched eigensolver (nrgk::diag_blocks): [[2,1,0],[1,2,0],[0,0,5]] has
    // eigenvalues 1, 3, 5; a 1 x 1 block is its own eigenvalue
    std::vector<nrgk::block> blocks(2);
    blocks[0] = nrgk::block{3, 3, {2, 1, 0, 1, 2, 0, 0, 0, 5}};
    blocks[1] = nrgk::block{1, 1, {-0.5}};
    const auto w = nrgk::diag_blocks(blocks);
    const double ev[3] = {1.0, 3.0, 5.0};
    bool ok = w.size() == 2 && w[0].size() == 3 && w[1].size() == 1 && std::fabs(w[1][0] + 0.5) < 1e-14;
    for (int i = 0; ok && i < 3; i++) ok = std::fabs(w[0][i] - ev[i]) < 1e-12;
    std::fprintf(stderr, "diag_blocks %s\n", ok ? "ok" : "wrong");
    if (!ok) return 4;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "%s\n", ex.what());
    return 1;
  }
  return 0;
}
