/* A D3Q19 duct written against OpenLB's API, for running the same program on
 * OpenLB's CPU_SISD platform and on GPU_HIP and comparing the results.
 *
 * BounceBack walls, a velocity inlet at x = 0 and a pressure outlet at
 * x = nx - 1 of one boundary family, chosen on the command line:
 *   (default)         boundary::LocalVelocity + boundary::LocalPressure
 *   --interpolated    boundary::InterpolatedVelocity + boundary::InterpolatedPressure
 * --smagorinsky runs a SmagorinskyBGKdynamics bulk (C = 0.14) next to the
 * boundaries' own BGK mixin, as aorta3d.cpp does.
 * --load FILE replaces the state after initialize() by OpenLB checkpoint FILE
 * (SuperLattice::load, then the switch to ProcessingContext::Simulation that
 * hands the loaded populations to the platform); --save FILE writes one after
 * the steps (switch to Evaluation, SuperLattice::save).
 * --cuboids N divides the domain into N cuboids (CuboidDecomposition with
 * OpenLB's default overlap of 3), whose blocks exchange their overlap and
 * whose boundary post-processors read cells of the neighbouring blocks.
 * The inflow velocity is uniform over the inlet face; the fluid starts at
 * rest.  After the steps the program writes the physical velocity and
 * pressure (SuperVTMwriter3D) and prints "nx, ny, nz, steps, MLUPs,
 * platform" (the Platform value block 0 runs on).
 */
#include <olb.h>

using namespace olb;
using namespace olb::descriptors;

using T = double;
#ifdef DUCT_FORCED
// built with -DDUCT_FORCED: a ForcedBGK bulk with a body force along the duct,
// next to boundaries with their BGK mixin (testFlow3dSolver.h:203-213)
using DESCRIPTOR = D3Q19<FORCE>;
using Bulk = ForcedBGKdynamics<T,DESCRIPTOR>;
#else
using DESCRIPTOR = D3Q19<>;
using Bulk = BGKdynamics<T,DESCRIPTOR>;
#endif
using Mixin = BGKdynamics<T,DESCRIPTOR>;

namespace {

void prepareGeometry(SuperGeometry<T,3>& geometry, IndicatorF3D<T>& box, T dx, T lx)
{
  geometry.rename(0, 2, box);
  geometry.rename(2, 1, {1, 1, 1});
  const T big = 1e3;
  // the x = 0 and x = nx - 1 wall layers next to the fluid: inlet and outlet
  IndicatorCuboid3D<T> inlet(Vector<T,3>(2 * dx, big, big), Vector<T,3>(-dx, -big / 2, -big / 2));
  IndicatorCuboid3D<T> outlet(Vector<T,3>(2 * dx, big, big), Vector<T,3>(lx - dx, -big / 2, -big / 2));
  geometry.rename(2, 3, 1, inlet);
  geometry.rename(2, 4, 1, outlet);
  geometry.clean(false);
  geometry.innerClean(false);
  geometry.checkForErrors(false);
}

template <typename INLET, typename OUTLET>
void prepareLattice(SuperLattice<T,DESCRIPTOR>& lattice, SuperGeometry<T,3>& geometry, T omega, bool smagorinsky)
{
  if (smagorinsky) {
    lattice.defineDynamics<SmagorinskyBGKdynamics<T,DESCRIPTOR>>(geometry, 1);
    lattice.setParameter<collision::LES::SMAGORINSKY>(T(0.14));
  } else {
    lattice.defineDynamics<Bulk>(geometry, 1);
  }
  boundary::set<boundary::BounceBack>(lattice, geometry, 2);
  boundary::set<INLET>(lattice, geometry, 3);
  boundary::set<OUTLET>(lattice, geometry, 4);
  lattice.setParameter<descriptors::OMEGA>(omega);
}

}

int main(int argc, char** argv)
{
  initialize(&argc, &argv, false, false);
  CLIreader args(argc, argv);
  const int nx = args.getValueOrFallback<int>("--nx", 40);
  const int ny = args.getValueOrFallback<int>("--ny", 14);
  const int nz = args.getValueOrFallback<int>("--nz", 12);
  const std::size_t steps = args.getValueOrFallback<std::size_t>("--steps", 200);
  const bool interpolated = args.contains("--interpolated");
  const int nCuboids = args.getValueOrFallback<int>("--cuboids", 1);
  const bool smagorinsky = args.contains("--smagorinsky");
  const std::string loadFrom = args.getValueOrFallback<std::string>("--load", "");
  const std::string saveTo = args.getValueOrFallback<std::string>("--save", "");
  singleton::directories().setOutputDir("./tmp/");

  // ny - 1 voxels per unit length, lattice inflow velocity 0.04, Re 20
  const int resolution = ny - 1;
  UnitConverterFromResolutionAndLatticeVelocity<T,DESCRIPTOR> const converter(
    resolution, 0.04, 1.0, 1.0, 1. / 20., 1.0);
  const T dx = converter.getPhysDeltaX();
  const T lx = T(nx - 1) * dx;
  IndicatorCuboid3D<T> box(Vector<T,3>(lx + 0.25 * dx, T(ny - 1) * dx + 0.25 * dx, T(nz - 1) * dx + 0.25 * dx),
                           Vector<T,3>(0, 0, 0));
  CuboidDecomposition3D<T> cuboids(box, dx, nCuboids);
  BlockLoadBalancer<T> load(0, 1, cuboids.size(), 0);
  SuperGeometry<T,3> geometry(cuboids, load);
  prepareGeometry(geometry, box, dx, lx);

  SuperLattice<T,DESCRIPTOR> lattice(cuboids, load, 3, converter);
  lattice.statisticsOff();
  const T omega = converter.getLatticeRelaxationFrequency();
  if (interpolated) {
    prepareLattice<boundary::InterpolatedVelocity<T,DESCRIPTOR,Mixin>,
                   boundary::InterpolatedPressure<T,DESCRIPTOR,Mixin>>(lattice, geometry, omega, smagorinsky);
  } else {
    prepareLattice<boundary::LocalVelocity<T,DESCRIPTOR,Mixin>,
                   boundary::LocalPressure<T,DESCRIPTOR,Mixin>>(lattice, geometry, omega, smagorinsky);
  }

#ifdef DUCT_FORCED
  AnalyticalConst3D<T,T> bodyForce(Vector<T,3>{T(2e-5), T(0), T(-5e-6)});
  lattice.defineField<FORCE>(geometry, 1, bodyForce);
#endif
  AnalyticalConst3D<T,T> rho(1);
  AnalyticalConst3D<T,T> rest(Vector<T,3>{});
  for (int m = 1; m <= 4; ++m) {
    lattice.iniEquilibrium(geometry, m, rho, rest);
    lattice.defineRhoU(geometry, m, rho, rest);
  }
  AnalyticalConst3D<T,T> inflow(Vector<T,3>{converter.getCharLatticeVelocity(), 0, 0});
  lattice.defineU(geometry, 3, inflow);
  lattice.initialize();
  if (!loadFrom.empty()) {
    if (!lattice.template load<false>(loadFrom)) {
      throw std::runtime_error("cannot load " + loadFrom);
    }
    lattice.setProcessingContext(ProcessingContext::Simulation);
  }

  util::Timer<T> timer(steps, geometry.getStatistics().getNvoxel());
  timer.start();
  for (std::size_t iT = 0; iT < steps; ++iT) {
    lattice.collideAndStream();
  }
  timer.stop();
  timer.update(steps);

  lattice.setProcessingContext(ProcessingContext::Evaluation);
  if (!saveTo.empty() && !lattice.template save<false>(saveTo)) {
    throw std::runtime_error("cannot save " + saveTo);
  }
  SuperVTMwriter3D<T> writer("duct3d");
  SuperLatticePhysVelocity3D<T,DESCRIPTOR> velocity(lattice, converter);
  SuperLatticePhysPressure3D<T,DESCRIPTOR> pressure(lattice, converter);
  writer.write(velocity);
  writer.write(pressure);
  std::cout << nx << ", " << ny << ", " << nz << ", " << steps << ", " << timer.getTotalMLUPs() << ", "
            << static_cast<int>(lattice.getBlock(0).getPlatform()) << std::endl;
  std::cout << "cuboids " << cuboids.size() << std::endl;
  return 0;
}
