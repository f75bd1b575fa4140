// The reference's examples/entropySIAM/main.cpp written against libnrgk's C++ API: the same calls in
// the same order (nrgcore, set_parameters, add_bath_site, update_internal_state, NrgData
// saveCurrentData / loadCurrentData, the backward loop that decrements nrg_iterations_cnt,
// fdmThermodynamics calcThermodynamics / saveFinalData, h5stream writes).  Changed: the includes, the
// model types (nrgk::site) and the parameters, which are those the reference's shipped
// resultSIAMU0.200000.h5 was produced with (eps = -U/2, V = sqrt(2 Gamma / pi) without the A_Lambda
// factor, K = 1024; the shipped source has eps = -0.4 U, the A_Lambda factor and K = 512).
// Usage: nrgk_entropy_siam <output.h5>
#include <cmath>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/nrgk/nrgcore.hpp"

using namespace nrgk;

double hopping(int site, double LAMBDA) {
  return 0.5 * (1.0 + 1.0 / LAMBDA) * (1. - std::pow(LAMBDA, -site - 1)) /
         std::sqrt((1.0 - std::pow(LAMBDA, -2. * site - 1)) * (1.0 - std::pow(LAMBDA, -2. * site - 3)));
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0] << " <output.h5>" << std::endl;
    return 2;
  }
  double LAMBDA = 2.50;
  int nMax{41};             // Number of NRG iteration
  size_t minIterations{0};  // minimum number of iteration
  double U_int = 0.20;
  double GAMMA = 0.0100;
  double V = std::sqrt(2.0 * GAMMA / std::acos(-1.));
  double eps = -U_int * 0.50;
  h5stream::h5stream rfile(argv[1]);
  site impurity = site::spinhalf(eps, U_int);
  // bath
  site bathModel = site::spinhalf(0, 0);
  nrgcore<site, site> siam(impurity, bathModel);
  siam.set_parameters(1024);        // set max number of states to be kept
  siam.add_bath_site({V, V}, 1.0);  // first site
  siam.update_internal_state();
  NrgData rawData(&siam);
  for (int in = 0; in < nMax; in++) {
    double rescale = 1.0;
    if (in > 0) {
      rescale = std::sqrt(LAMBDA);
    }
    siam.add_bath_site({hopping(in, LAMBDA), hopping(in, LAMBDA)}, rescale);
    if (siam.checkHigherEnergyDiscarded()) {
      rawData.saveCurrentData();
    } else {
      minIterations++;
    }
    siam.update_internal_state();
    rfile.write(siam.all_eigenvalue,  // Save the eigenvalue of the current iteration
                "Eigenvalues" + std::to_string(in));
  }
  // Backward iteration
  double vMax = 2;  // Max Voltage can be 2* bandwidth
  double vMin = 1e-10;
  int tPoints = 200;
  double delTime = (tPoints - 1.0) / (std::log(vMax / vMin));
  std::vector<double> temperatureArray(tPoints);
  for (int i = 0; i < tPoints; i++) {
    temperatureArray[i] = vMin * std::exp(std::fabs(i) * 1. / delTime);
  }
  fdmThermodynamics fdmTc(&siam, temperatureArray);
  for (size_t in = nMax; in > minIterations; in--) {
    double enScale = std::pow(LAMBDA, -(1.0 * static_cast<double>(in) - 1.0) / 2.0);
    siam.nrg_iterations_cnt--;
    rawData.loadCurrentData();
    fdmTc.calcThermodynamics(enScale);
  }
  fdmTc.saveFinalData(&rfile);
  rfile.close();
  std::cout << "entropySIAM done: " << nMax << " sites, " << (nMax - minIterations) << " shells" << std::endl;
  return 0;
}
