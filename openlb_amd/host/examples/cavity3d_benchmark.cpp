// Lid-driven cavity benchmark on the MI355X backend, written against the C++
// host layer the same way OpenLB's harness is written against SuperLattice
// (examples/laminar/cavity3dBenchmark/cavity3d.cpp:154-250):
//   * resolution `size`: (size+1)^3 cells, every one a voxel with dynamics;
//   * by default OpenLB's LID_DRIVEN configuration (:32-36, 68-76): LocalVelocity
//     boundaries on the whole box, at rest on the walls and moving along x at
//     the lattice velocity 0.01 on the lid (the top face y = size without its
//     rim, :39-57); with --bounce-back the LID_DRIVEN_BOUNCE_BACK variant
//     (:70-71): boundary cells BounceBack, the whole top face a
//     BounceBackVelocity lid;
//   * BGK bulk, tau from UnitConverterFromResolutionAndLatticeVelocity with
//     physical viscosity 1e-3, velocity 1, length 1 (nu_lattice = 1e-5 size);
//   * 10 warm-up steps, then `steps` timed steps bracketed by device syncs;
//   * prints "size, steps, cuboids, ranks, threads, MLUPs" like the original.
// Extra flags: --fp64 (double storage), --dump <file> (populations after the
// run, double [q][nx][ny][nz]), --tau (override), --stats (lattice statistics
// after the run), --load <name> (resume from OpenLB checkpoint
// <name>_rank0_size1.dat right after initialize(), as lattice.load(name)),
// --save <name> (write that checkpoint after the run, as lattice.save(name)),
// --vtk <dir> (after the run, SuperVTMwriter3D "cavity3d" into <dir> with
// physVelocity and physPressure, as cavity3d.cpp:90-111 writes them, through
// UnitConverter(1/size, 0.01/size, 1, 1, 1e-3, 1)).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "olb_hip.h"

using namespace olb::hip;

template <typename T>
int run(int size, long steps, double tau, const char* dump, bool stats, bool bounceBack, const char* load,
        const char* save, const char* vtk)
{
  using DESCRIPTOR = descriptors::D3Q19;
  const int n = size + 1;
  SuperLattice<T, DESCRIPTOR> lattice({n, n, n}, {false, false, false});
  lattice.statisticsOff();

  auto boundary = [n](int x, int y, int z) {
    return x == 0 || y == 0 || z == 0 || x == n - 1 || y == n - 1 || z == n - 1;
  };
  using BulkDynamics = BGKdynamics<T, DESCRIPTOR>;
  lattice.template defineDynamics<BulkDynamics>([&](int x, int y, int z) { return !boundary(x, y, z); });
  if (bounceBack) {
    lattice.template defineDynamics<BounceBack<T, DESCRIPTOR>>(boundary);
    lattice.template defineDynamics<BounceBackVelocity<T, DESCRIPTOR>>([n](int, int y, int) { return y == n - 1; });
    lattice.defineWall(T(1), {T(0.01), T(0), T(0)});
  } else {
    auto lid = [n](int x, int y, int z) { return y == n - 1 && x > 0 && x < n - 1 && z > 0 && z < n - 1; };
    boundary::set<boundary::LocalVelocity<T, DESCRIPTOR, BulkDynamics>>(
      lattice, [&](int x, int y, int z) { return boundary(x, y, z) && !lid(x, y, z); });
    lattice.defineWall(T(1), {T(0), T(0), T(0)});
    boundary::set<boundary::LocalVelocity<T, DESCRIPTOR, BulkDynamics>>(lattice, lid);
    lattice.defineWall(T(1), {T(0.01), T(0), T(0)});
  }
  lattice.template setParameter<descriptors::OMEGA>(T(1.0 / tau));

  lattice.iniEquilibrium([](int, int, int) { return true; }, [](int, int, int) { return 1.0; },
                         [](int, int, int) { return std::array<double, 3>{0.0, 0.0, 0.0}; });
  lattice.initialize();
  if (load && !lattice.template load<false>(load)) throw std::runtime_error(std::string("cannot read checkpoint ") + load);

  for (int i = 0; i < 10; ++i) lattice.collideAndStream();
  lattice.synchronize();
  const auto t0 = std::chrono::steady_clock::now();
  for (long i = 0; i < steps; ++i) lattice.collideAndStream();
  lattice.synchronize();
  const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const double mlups = double(lattice.ncore()) * double(steps) / seconds / 1e6;

  lattice.setProcessingContext(ProcessingContext::Evaluation);
  if (save && !lattice.template save<false>(save)) throw std::runtime_error(std::string("cannot write checkpoint ") + save);
  if (vtk) {
    // UnitConverterFromResolutionAndLatticeVelocity(N, 0.01, 1, 1, 1e-3, 1): dx = 1/N, dt = 0.01 dx
    const T dx = T(1) / T(size), dt = T(0.01) * dx;
    const UnitConverter<T, DESCRIPTOR> converter(dx, dt, T(1), T(1), T(1e-3), T(1));
    SuperVTMwriter3D<T>::setVtkOutDir(std::string(vtk) + "/");
    SuperVTMwriter3D<T> writer("cavity3d");
    writer.setDeltaR(converter.getPhysDeltaX());
    SuperLatticePhysVelocity3D<T, DESCRIPTOR> velocity(lattice, converter);
    SuperLatticePhysPressure3D<T, DESCRIPTOR> pressure(lattice, converter);
    writer.addFunctor(velocity);
    writer.addFunctor(pressure);
    writer.createMasterFile();
    writer.write(int(10 + steps));
    std::printf("vtk %.17g %.17g %.17g %.17g\n", double(converter.getConversionFactorVelocity()),
                double(converter.getConversionFactorPressure()), double(converter.getCharPhysPressure()),
                double(converter.getPhysDeltaX()));
    std::fflush(stdout);
  }
  std::cout << size << ", " << steps << ", " << 1 << ", " << 1 << ", " << 1 << ", " << mlups << std::endl;
  if (stats) {
    const auto st = lattice.getStatistics();
    std::printf("statistics %zu %.17g %.17g %.17g\n", st.getNumCells(), st.getAverageRho(), st.getAverageEnergy(),
                st.getMaxU());
    std::fflush(stdout);
  }
  if (dump) {
    std::FILE* f = std::fopen(dump, "wb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + dump);
    const auto& pop = lattice.populations();
    std::fwrite(pop.data(), sizeof(double), pop.size(), f);
    std::fclose(f);
  }
  return 0;
}

int main(int argc, char** argv)
{
  int size = 100;
  long steps = 100;
  bool fp64 = false;
  const char* dump = nullptr;
  bool stats = false;
  bool bounceBack = false;
  const char* load = nullptr;
  const char* save = nullptr;
  const char* vtk = nullptr;
  double tau = -1;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--size") && i + 1 < argc) size = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--steps") && i + 1 < argc) steps = std::atol(argv[++i]);
    else if (!std::strcmp(argv[i], "--fp64")) fp64 = true;
    else if (!std::strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
    else if (!std::strcmp(argv[i], "--tau") && i + 1 < argc) tau = std::atof(argv[++i]);
    else if (!std::strcmp(argv[i], "--stats")) stats = true;
    else if (!std::strcmp(argv[i], "--bounce-back")) bounceBack = true;
    else if (!std::strcmp(argv[i], "--load") && i + 1 < argc) load = argv[++i];
    else if (!std::strcmp(argv[i], "--save") && i + 1 < argc) save = argv[++i];
    else if (!std::strcmp(argv[i], "--vtk") && i + 1 < argc) vtk = argv[++i];
    else if (!std::strcmp(argv[i], "--no-results")) continue;
    else {
      std::fprintf(stderr, "usage: %s [--size N] [--steps K] [--fp64] [--tau T] [--dump FILE] [--stats] [--bounce-back] [--load NAME] [--save NAME] [--vtk DIR]\n", argv[0]);
      return 2;
    }
  }
  if (tau <= 0) tau = 0.5 + 3.0 * 1e-5 * size;  // nu = nu_phys * dt / dx^2 with u_lattice 0.01
  try {
    return fp64 ? run<double>(size, steps, tau, dump, stats, bounceBack, load, save, vtk)
                : run<float>(size, steps, tau, dump, stats, bounceBack, load, save, vtk);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
