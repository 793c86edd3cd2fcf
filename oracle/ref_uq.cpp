// Sampling-point and statistics vectors of OpenLB's uncertainty-quantification
// module (src/uq, header-only), for the fixtures under tests/golden/.
//
// Test infrastructure: built from the reference sources in place by
// oracle/Makefile into oracle/_ref/olb_ref_uq and run in the build container
// only (tests/golden/gen_golden.py --uq-only).  Usage: olb_ref_uq <sobol
// direction-number file>.  Prints one JSON object: for every configuration the
// sampling points that UncertaintyQuantification<double>::getSamplingPoints()
// returns and the mean / std it reports for the test response
// g(x) = sum_j (j + 1) x_j^2 + sin(x_0) evaluated at those points
// (uncertaintyQuantification.hh:180-330).
#include <stdexcept>

#include "uq/uncertaintyQuantification.h"
#include "uq/generalizedPolynomialChaos.hh"
#include "uq/polynomial.hh"
#include "uq/uncertaintyQuantification.hh"
#include "uq/filesIO.hh"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

using namespace olb::uq;
using T = double;

static T response(const std::vector<T>& x)
{
  T g = std::sin(x[0]);
  for (std::size_t j = 0; j < x.size(); ++j) g += T(j + 1) * x[j] * x[j];
  return g;
}

static void emit(const char* name, UncertaintyQuantification<T>& uq, bool& first)
{
  auto pts = uq.getSamplingPoints();
  std::vector<T> g;
  for (const auto& p : pts) g.push_back(response(p));
  std::printf("%s\"%s\": {\"points\": [", first ? "" : ", ", name);
  for (std::size_t i = 0; i < pts.size(); ++i) {
    std::printf("%s[", i ? ", " : "");
    for (std::size_t j = 0; j < pts[i].size(); ++j) std::printf("%s%.17g", j ? ", " : "", pts[i][j]);
    std::printf("]");
  }
  std::printf("], \"mean\": %.17g, \"std\": %.17g}", uq.mean(g), uq.std(g));
  first = false;
}

int main(int argc, char** argv)
{
  const std::string dir = argc > 1 ? argv[1] : "new-joe-kuo-6.21201";
  bool first = true;
  std::printf("{");
  {
    UncertaintyQuantification<T> uq(UQMethod::MonteCarlo);
    uq.initializeMonteCarlo(10, uniform(T(0.9), T(1.1)), 123456u);
    emit("mc_uniform_1d", uq, first);
  }
  {
    UncertaintyQuantification<T> uq(UQMethod::MonteCarlo);
    uq.initializeMonteCarlo(17, normal(T(0.05), T(0.01)), 7u);
    emit("mc_normal_1d", uq, first);
  }
  {
    UncertaintyQuantification<T> uq(UQMethod::MonteCarlo);
    uq.initializeMonteCarlo(12, joint<T>({uniform(T(0.52), T(1.2)), uniform(T(0.02), T(0.1)), normal(T(0.05), T(0.01))}),
                            20261204u);
    emit("mc_joint_3d", uq, first);
  }
  {
    UncertaintyQuantification<T> uq(UQMethod::GPC);
    uq.initializeGPC(2, 5, uniform(T(0.9), T(1.1)));
    emit("gpc_uniform_o2_nq5", uq, first);
  }
  {
    UncertaintyQuantification<T> uq(UQMethod::GPC);
    uq.initializeGPC(3, 4, normal(T(0.05), T(0.01)));
    emit("gpc_normal_o3_nq4", uq, first);
  }
  {
    UncertaintyQuantification<T> uq(UQMethod::GPC);
    uq.initializeGPC(2, 3, std::vector<Distribution<T>>{uniform(T(0.52), T(1.2)), normal(T(0.05), T(0.01))});
    emit("gpc_joint_o2_nq3", uq, first);
  }
  {
    UncertaintyQuantification<T> uq(UQMethod::QuasiMonteCarlo);
    uq.initializeQuasiMonteCarlo(16, std::vector<Distribution<T>>{uniform(T(0.52), T(1.2)), uniform(T(0.02), T(0.1))},
                                 dir, GeneratorType::Sobol);
    emit("qmc_sobol_2d", uq, first);
  }
  {
    UncertaintyQuantification<T> uq(UQMethod::QuasiMonteCarlo);
    uq.initializeQuasiMonteCarlo(11, std::vector<Distribution<T>>{uniform(T(0.9), T(1.1)), uniform(T(-1), T(1)),
                                                                  uniform(T(0), T(2))},
                                 dir, GeneratorType::Halton);
    emit("qmc_halton_3d", uq, first);
  }
  std::printf("}\n");
  return 0;
}
