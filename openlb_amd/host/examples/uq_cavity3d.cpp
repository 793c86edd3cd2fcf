// Uncertainty quantification of the lid-driven cavity on the ensemble: the
// study of examples/uncertaintyQuantification/cavity3d/cavity3d.cpp:97-154
// (lid velocity uniform in [0.9 U, 1.1 U], stochastic collocation with
// generalized polynomial chaos of order 2 on 5 Gauss points, or Monte Carlo),
// written against the C++ host layer.  OpenLB simulates one sample after the
// other; here every sample is set up as a SuperLattice, added to one
// Ensemble, and all of them advance in the same launches.  Afterwards the
// physVelocity of every member is exported on the device and the mean and
// standard deviation are taken point by point as OpenLB's UQ module defines
// them (src/uq/generalizedPolynomialChaos.hh randomToChaos / mean / std;
// uncertaintyQuantification.hh:249-330 for Monte Carlo).
//
// Flags: --size N (cavity of (N+1)^3 cells), --steps K, --mc M --seed S
// (Monte Carlo with M samples instead of GPC), --vtk DIR (mean and std as a
// VTK piece), --dump FILE (member velocities as float32, then mean and std as
// float64).  Prints the sampling points and a summary line.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "olb_hip.h"

using namespace olb::hip;
using T = float;
using DESCRIPTOR = descriptors::D3Q19;

// Gauss points and probability weights of the uniform measure on [-1, 1]:
// roots of the Legendre polynomial P_n by Newton's method from the Chebyshev
// guesses, w = 1 / ((1 - x^2) P_n'(x)^2) (halved so that the weights sum to 1).
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w)
{
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = std::acos(-1.0);
  for (int k = 0; k < n; ++k) {
    double r = std::cos(pi * (k + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = r;
      for (int j = 2; j <= n; ++j) {
        const double p2 = ((2.0 * j - 1.0) * r * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      const double pn = n == 1 ? r : p1;
      const double pnm1 = n == 1 ? 1.0 : p0;
      dp = n * (r * pn - pnm1) / (r * r - 1.0);
      const double step = pn / dp;
      r -= step;
      if (std::fabs(step) < 1e-16) break;
    }
    x[n - 1 - k] = r;
    w[n - 1 - k] = 1.0 / ((1.0 - r * r) * dp * dp);
  }
}

static double legendre(int k, double x)
{
  double p0 = 1.0, p1 = x;
  if (k == 0) return 1.0;
  for (int j = 2; j <= k; ++j) {
    const double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

// VTK functor of a field computed on the host (mean or std)
struct HostField {
  std::string name;
  int dim;
  std::array<int, 3> n;
  std::vector<float> data;
  std::string& getName() { return name; }
  int getTargetDim() const { return dim; }
  const std::array<int, 3>& extent() const { return n; }
  std::vector<float> evaluate(int) { return data; }
};

int main(int argc, char** argv)
{
  int size = 30;
  long steps = 2000;
  int mc = 0;
  unsigned seed = 123456u;
  const char* vtk = nullptr;
  const char* dump = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--size") && i + 1 < argc) size = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--steps") && i + 1 < argc) steps = std::atol(argv[++i]);
    else if (!std::strcmp(argv[i], "--mc") && i + 1 < argc) mc = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) seed = unsigned(std::atol(argv[++i]));
    else if (!std::strcmp(argv[i], "--vtk") && i + 1 < argc) vtk = argv[++i];
    else if (!std::strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
    else {
      std::fprintf(stderr, "usage: %s [--size N] [--steps K] [--mc M [--seed S]] [--vtk DIR] [--dump FILE]\n", argv[0]);
      return 2;
    }
  }
  try {
    const int n = size + 1;
    const T latticeU = T(0.01);
    const double lo = 0.9, hi = 1.1;   // lid speed factor, uniform (cavity3d.cpp:104)
    const double tau = 0.5 + 3.0 * 1e-5 * size;
    // sampling points (getSamplingPoints) and, for GPC, the collocation weights
    std::vector<double> samples, xi, wq;
    const int order = 2;
    if (mc > 0) {
      std::mt19937 rng(seed);
      std::uniform_real_distribution<double> dist(lo, hi);   // MonteCarlo::generateSamples, monteCarlo.h:60-110
      for (int k = 0; k < mc; ++k) samples.push_back(dist(rng));
    } else {
      gauss_legendre(5, xi, wq);
      for (double x : xi) samples.push_back(0.5 * (hi - lo) * x + 0.5 * (lo + hi));
    }
    std::printf("samples");
    for (double v : samples) std::printf(" %.17g", v);
    std::printf("\ntau %.17g\n", tau);
    std::fflush(stdout);

    // one member per sample, set up as OpenLB's cavity (cavity3d.cpp:39-76, LocalVelocity)
    SuperLattice<T, DESCRIPTOR> probe({n, n, n}, {false, false, false});
    Ensemble<T, DESCRIPTOR> ensemble(int(samples.size()),
                                     (probe.q * std::size_t(n + 2) * (n + 2) * (n + 2) * sizeof(T) + (1 << 16)) *
                                         samples.size(),
                                     std::size_t(n) * n * n * samples.size() + (1 << 16), LBMH_BGK);
    std::vector<int> members;
    for (double factor : samples) {
      SuperLattice<T, DESCRIPTOR> lattice({n, n, n}, {false, false, false});
      auto boundary = [n](int x, int y, int z) {
        return x == 0 || y == 0 || z == 0 || x == n - 1 || y == n - 1 || z == n - 1;
      };
      auto lid = [n](int x, int y, int z) { return y == n - 1 && x > 0 && x < n - 1 && z > 0 && z < n - 1; };
      using Bulk = BGKdynamics<T, DESCRIPTOR>;
      lattice.template defineDynamics<Bulk>([&](int x, int y, int z) { return !boundary(x, y, z); });
      boundary::set<boundary::LocalVelocity<T, DESCRIPTOR, Bulk>>(
          lattice, [&](int x, int y, int z) { return boundary(x, y, z) && !lid(x, y, z); });
      lattice.defineWall(T(1), {T(0), T(0), T(0)});
      boundary::set<boundary::LocalVelocity<T, DESCRIPTOR, Bulk>>(lattice, lid);
      lattice.defineWall(T(1), {T(latticeU * T(factor)), T(0), T(0)});
      lattice.template setParameter<descriptors::OMEGA>(T(1.0 / tau));
      lattice.iniEquilibrium([](int, int, int) { return true; }, [](int, int, int) { return 1.0; },
                             [](int, int, int) { return std::array<double, 3>{0.0, 0.0, 0.0}; });
      members.push_back(ensemble.add(lattice, steps));
    }
    ensemble.collideAndStream(steps);
    for (int m : members) {
      const auto st = ensemble.status(m);
      if (st.first != steps) throw std::runtime_error("member did not finish its steps");
    }

    // physVelocity of every member, then mean and std point by point
    const UnitConverter<T, DESCRIPTOR> converter(T(1) / T(size), T(0.01) / T(size), T(1), T(1), T(1e-3), T(1));
    std::vector<std::vector<float>> fields;
    for (int m : members) {
      SuperLatticePhysVelocity3D<T, DESCRIPTOR> f(ensemble, m, converter);
      fields.push_back(f.evaluate(1));
    }
    const std::size_t P = fields.front().size();
    const std::size_t S = fields.size();
    std::vector<double> mean(P, 0.0), std_(P, 0.0);
    if (mc > 0) {
      for (std::size_t p = 0; p < P; ++p) {
        double s = 0.0;
        for (std::size_t k = 0; k < S; ++k) s += fields[k][p];
        mean[p] = s / double(S);
        double q = 0.0;
        for (std::size_t k = 0; k < S; ++k) q += (fields[k][p] - mean[p]) * (fields[k][p] - mean[p]);
        std_[p] = std::sqrt(q / double(S - 1));
      }
    } else {
      // chaos coefficients c_j = sum_k w_k f_k P_j(xi_k) / <P_j^2>, mean c_0, var sum_{j>0} <P_j^2> c_j^2
      std::vector<double> t2(order + 1, 0.0);
      for (int j = 0; j <= order; ++j)
        for (std::size_t k = 0; k < S; ++k) t2[j] += legendre(j, xi[k]) * legendre(j, xi[k]) * wq[k];
      for (std::size_t p = 0; p < P; ++p) {
        double var = 0.0;
        for (int j = 0; j <= order; ++j) {
          double c = 0.0;
          for (std::size_t k = 0; k < S; ++k) c += wq[k] * fields[k][p] * legendre(j, xi[k]);
          c /= t2[j];
          if (j == 0) mean[p] = c;
          else var += t2[j] * c * c;
        }
        std_[p] = std::sqrt(var);
      }
    }
    double maxMean = 0.0, maxStd = 0.0;
    for (std::size_t p = 0; p < P; ++p) {
      maxMean = std::max(maxMean, std::fabs(mean[p]));
      maxStd = std::max(maxStd, std_[p]);
    }
    std::printf("uq %s members %zu steps %ld max|mean u| %.6g max std %.6g\n", mc > 0 ? "monte_carlo" : "gpc", S, steps,
                maxMean, maxStd);
    std::fflush(stdout);
    if (vtk) {
      SuperVTMwriter3D<T>::setVtkOutDir(std::string(vtk) + "/");
      SuperVTMwriter3D<T> writer("uq_cavity3d");
      writer.setDeltaR(converter.getPhysDeltaX());
      HostField fm{"physVelocityMean", 3, {n, n, n}, std::vector<float>(mean.begin(), mean.end())};
      HostField fs{"physVelocityStd", 3, {n, n, n}, std::vector<float>(std_.begin(), std_.end())};
      writer.addFunctor(fm);
      writer.addFunctor(fs);
      writer.createMasterFile();
      writer.write(int(steps));
    }
    if (dump) {
      std::FILE* f = std::fopen(dump, "wb");
      if (!f) throw std::runtime_error(std::string("cannot open ") + dump);
      for (const auto& a : fields) std::fwrite(a.data(), sizeof(float), a.size(), f);
      std::fwrite(mean.data(), sizeof(double), mean.size(), f);
      std::fwrite(std_.data(), sizeof(double), std_.size(), f);
      std::fclose(f);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
