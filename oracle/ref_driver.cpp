// Golden-vector driver for the compiled OpenLB reference (CPU_SISD platform).
//
// This program is test infrastructure: it links against the OpenLB sources in
// the reference tree (built by oracle/Makefile into oracle/_ref/) and is only
// run in the build container to produce the fixtures under tests/golden/.
// Nothing that ships to a GPU runs it.
//
// Input: a case directory holding
//   case.txt  key/value lines (q, dyn, nx, ny, nz, px, py, pz, omega, magic,
//             smago, fx, fy, fz, steps, wall <id> <rho> <ux> <uy> <uz>,
//             local <id> velocity|pressure|interp_velocity|interp_pressure,
//             split <width>: CuboidDecomposition::splitRegular(0, width),
//             i.e. max(1, n_d / width) cuboids per axis cut by Cuboid::divide,
//             cuboidDecomposition.hh:755-763, cuboid.hh:184-211)
//             overlap <n>: the geometry's (and lattice's) overlap, default 1;
//             with this key the overlap cells inside the domain (or its
//             periodic images) are initialised like the core cells, as an
//             example program's defineRhoU / iniEquilibrium over a material
//             indicator does (blockLattice.hh:69-120 iterate every spatial
//             location, overlap included).  OpenLB's own default is 3
//             (superGeometry.h:94-96).
//   ids.bin   uint8  [nx][ny][nz]  per-cell dynamics id
//             (0 none, 1 bulk, 2 BounceBack, 3.. BounceBackVelocity walls)
//   rhou.bin  double [4][nx][ny][nz] initial rho, ux, uy, uz
//   force.bin (optional) double [3][nx][ny][nz] per-cell FORCE of the bulk
//             cells, in place of the uniform fx fy fz
// Output (same directory):
//   pops_<step>.bin  double [q][nx][ny][nz] shifted populations of the core
//                    cells after <step> collideAndStream() calls (0 = after
//                    initialize()).
//   checkpoint_<step>_rank0_size1.dat  (only with OLB_REF_SAVE set): OpenLB's
//                    own serialization of the lattice after <step> steps
//                    (Serializable::save, serializer.hh:81-97, base64 text of
//                    the blocks of superLattice.hh:806-834).
//   With OLB_REF_LOAD=<path> (a checkpoint written by either side, without
//   its "_rank0_size1.dat" suffix) the state after initialize() is replaced by
//   that checkpoint (Serializable::load, then postLoad) and <step> counts the
//   steps taken after the load.
//   vtkData/fields.pvd, vtkData/data/fields_iT<step>.vtm and
//   vtkData/data/fields_iT<step>iC00000.vti  (only with OLB_REF_VTK set): after
//                    the last step, SuperVTMwriter3D<T> "fields" with the
//                    functors density, velocity, physVelocity and physPressure
//                    (io/superVtmWriter3D.hh; functors/lattice/), the physical
//                    ones through UnitConverter(dx 0.01, dt 1e-4, L 0.1, U 1,
//                    nu 1e-5, rho 1000, p0 100).
//   normals.bin      int8 [4][nx][ny][nz] (only with OLB_REF_NORMALS set):
//                    type and discrete normal that OpenLB's
//                    computeBoundaryTypeAndNormal gives every core cell that
//                    is neither fluid (material 1) nor outside (material 0);
//                    type -1 elsewhere.
//
// A material listed as `local <id> velocity` (`pressure`) gets OpenLB's
// LocalVelocity (LocalPressure) boundary with BGKdynamics mixed in -- the
// bulk dynamics of cavity3d.cpp:72-73, and the boundaries' default mixin next
// to any other bulk (aorta3d.cpp:118-127); `interp_velocity`
// (`interp_pressure`) gets InterpolatedVelocity (InterpolatedPressure).  Its
// `wall` line gives the velocity (density) that defineRhoU assigns.
//
// The setup follows the reference's example programs:
//   geometry/materials  examples/laminar/cavity3dBenchmark/cavity3d.cpp:39-57
//   dynamics + params   cavity3d.cpp:59-75, tgv3d.cpp:240-253
//   init                tgv3d.cpp:313-326 (defineRhoU then iniEquilibrium)
//   time loop           src/core/superLattice.hh:601-663
#include "olb.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

using namespace olb;

#ifndef REF_REAL
#define REF_REAL double
#endif
using T = REF_REAL;

struct WallSpec { int id; double rho, u[3]; };

struct CaseSpec {
  int q = 19;
  std::string dyn = "bgk";
  int n[3] = {8, 8, 8};
  int periodic[3] = {1, 1, 1};
  double omega = 1.0, magic = 0.25, smago = 0.1;
  double force[3] = {0, 0, 0};
  std::vector<long> steps;
  std::vector<WallSpec> walls;
  std::map<int, std::string> local;
  int split = 0;   // splitRegular width; 0 keeps one cuboid
  int overlap = 0; // 0: overlap 1 with core-only initialisation
  std::vector<unsigned char> ids;
  std::vector<double> rhou;
  std::vector<double> forceField;  // [3][nCore], empty: the uniform force
};

static std::size_t coreIndex(const CaseSpec& c, int x, int y, int z)
{
  return (std::size_t(x) * c.n[1] + y) * c.n[2] + z;
}

static CaseSpec readCase(const std::string& dir)
{
  CaseSpec c;
  std::ifstream in(dir + "/case.txt");
  if (!in) throw std::runtime_error("cannot open case.txt in " + dir);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key)) continue;
    if (key == "q") ls >> c.q;
    else if (key == "dyn") ls >> c.dyn;
    else if (key == "nx") ls >> c.n[0];
    else if (key == "ny") ls >> c.n[1];
    else if (key == "nz") ls >> c.n[2];
    else if (key == "px") ls >> c.periodic[0];
    else if (key == "py") ls >> c.periodic[1];
    else if (key == "pz") ls >> c.periodic[2];
    else if (key == "omega") ls >> c.omega;
    else if (key == "magic") ls >> c.magic;
    else if (key == "smago") ls >> c.smago;
    else if (key == "fx") ls >> c.force[0];
    else if (key == "fy") ls >> c.force[1];
    else if (key == "fz") ls >> c.force[2];
    else if (key == "steps") { long s; while (ls >> s) c.steps.push_back(s); }
    else if (key == "split") ls >> c.split;
    else if (key == "overlap") ls >> c.overlap;
    else if (key == "local") {
      int id; std::string kind; ls >> id >> kind;
      if (kind != "velocity" && kind != "pressure" && kind != "interp_velocity" && kind != "interp_pressure")
        throw std::runtime_error("local kind " + kind);
      c.local[id] = kind;
    }
    else if (key == "wall") {
      WallSpec w; ls >> w.id >> w.rho >> w.u[0] >> w.u[1] >> w.u[2];
      c.walls.push_back(w);
    }
    else throw std::runtime_error("unknown key " + key);
  }
  const std::size_t nCore = std::size_t(c.n[0]) * c.n[1] * c.n[2];
  c.ids.resize(nCore);
  c.rhou.resize(4 * nCore);
  std::FILE* f = std::fopen((dir + "/ids.bin").c_str(), "rb");
  if (!f || std::fread(c.ids.data(), 1, nCore, f) != nCore) throw std::runtime_error("ids.bin");
  std::fclose(f);
  f = std::fopen((dir + "/rhou.bin").c_str(), "rb");
  if (!f || std::fread(c.rhou.data(), sizeof(double), 4 * nCore, f) != 4 * nCore) throw std::runtime_error("rhou.bin");
  std::fclose(f);
  f = std::fopen((dir + "/force.bin").c_str(), "rb");
  if (f) {
    c.forceField.resize(3 * nCore);
    if (std::fread(c.forceField.data(), sizeof(double), 3 * nCore, f) != 3 * nCore) throw std::runtime_error("force.bin");
    std::fclose(f);
  }
  return c;
}

// Calls f(local block index, local cell, global cell) for every core cell of
// every block this process holds; cuboid origins are integers (spacing 1,
// origin 0).
template <typename F>
static void forEachCoreCell(SuperGeometry<T,3>& geometry, F f)
{
  auto& load = geometry.getLoadBalancer();
  auto& cuboids = geometry.getCuboidDecomposition();
  for (int iC = 0; iC < load.size(); ++iC) {
    const auto& cuboid = cuboids.get(load.glob(iC));
    const auto origin = cuboid.getOrigin();
    const auto extent = cuboid.getExtent();
    const int o[3] = {int(std::lround(double(origin[0]))), int(std::lround(double(origin[1]))),
                      int(std::lround(double(origin[2])))};
    for (int x = 0; x < extent[0]; ++x)
      for (int y = 0; y < extent[1]; ++y)
        for (int z = 0; z < extent[2]; ++z)
          f(iC, LatticeR<3>{x, y, z}, LatticeR<3>{o[0] + x, o[1] + y, o[2] + z});
  }
}

// Calls f(local block index, local cell, global cell) for every spatial
// location of every block -- core and `overlap` layers -- that lies in the
// domain or in one of its periodic images (global cell wrapped into the
// domain).
template <typename F>
static void forEachSpatialCell(const CaseSpec& c, SuperGeometry<T,3>& geometry, int overlap, F f)
{
  auto& load = geometry.getLoadBalancer();
  auto& cuboids = geometry.getCuboidDecomposition();
  for (int iC = 0; iC < load.size(); ++iC) {
    const auto& cuboid = cuboids.get(load.glob(iC));
    const auto origin = cuboid.getOrigin();
    const auto extent = cuboid.getExtent();
    const int o[3] = {int(std::lround(double(origin[0]))), int(std::lround(double(origin[1]))),
                      int(std::lround(double(origin[2])))};
    for (int x = -overlap; x < extent[0] + overlap; ++x)
      for (int y = -overlap; y < extent[1] + overlap; ++y)
        for (int z = -overlap; z < extent[2] + overlap; ++z) {
          int g[3] = {o[0] + x, o[1] + y, o[2] + z};
          bool inside = true;
          for (int d = 0; d < 3; ++d) {
            if (g[d] < 0 || g[d] >= c.n[d]) {
              if (!c.periodic[d]) inside = false;
              else g[d] = (g[d] % c.n[d] + c.n[d]) % c.n[d];
            }
          }
          if (inside) f(iC, LatticeR<3>{x, y, z}, LatticeR<3>{g[0], g[1], g[2]});
        }
  }
}

template <typename DESCRIPTOR>
static void dumpPopulations(const CaseSpec& c, SuperGeometry<T,3>& geometry, SuperLattice<T,DESCRIPTOR>& lattice,
                            const std::string& dir, long step)
{
  const std::size_t nCore = std::size_t(c.n[0]) * c.n[1] * c.n[2];
  std::vector<double> out(std::size_t(DESCRIPTOR::q) * nCore);
  forEachCoreCell(geometry, [&](int iC, LatticeR<3> local, LatticeR<3> global) {
    auto cell = lattice.getBlock(iC).get(local);
    const std::size_t k = coreIndex(c, global[0], global[1], global[2]);
    for (int i = 0; i < int(DESCRIPTOR::q); ++i) {
      out[std::size_t(i) * nCore + k] = double(cell[i]);
    }
  });
  const std::string name = dir + "/pops_" + std::to_string(step) + ".bin";
  std::FILE* f = std::fopen(name.c_str(), "wb");
  if (!f || std::fwrite(out.data(), sizeof(double), out.size(), f) != out.size()) {
    throw std::runtime_error("cannot write " + name);
  }
  std::fclose(f);
}

// Boundary type and discrete normal of every non-fluid, non-outside core
// cell, as OpenLB's boundary setters compute them (setBoundary.h:83-101).
static void dumpNormals(const CaseSpec& c, SuperGeometry<T,3>& geometry, const std::string& dir)
{
  auto fluid = geometry.getMaterialIndicator(1);
  auto outside = geometry.getMaterialIndicator(0);
  auto& fluidI = fluid->getBlockIndicatorF(0);
  auto& outsideI = outside->getBlockIndicatorF(0);
  const std::size_t nCore = c.ids.size();
  std::vector<signed char> out(4 * nCore, 0);
  for (int x = 0; x < c.n[0]; ++x)
    for (int y = 0; y < c.n[1]; ++y)
      for (int z = 0; z < c.n[2]; ++z) {
        const std::size_t k = coreIndex(c, x, y, z);
        out[k] = -1;
        if (c.ids[k] == 0 || c.ids[k] == 1) continue;
        const auto [type, normal] = computeBoundaryTypeAndNormal(fluidI, outsideI, Vector<int,3>{x, y, z});
        out[k] = static_cast<signed char>(static_cast<int>(type));
        for (int d = 0; d < 3; ++d) out[(d + 1) * nCore + k] = static_cast<signed char>(normal[d]);
      }
  const std::string name = dir + "/normals.bin";
  std::FILE* f = std::fopen(name.c_str(), "wb");
  if (!f || std::fwrite(out.data(), 1, out.size(), f) != out.size()) throw std::runtime_error("cannot write " + name);
  std::fclose(f);
}

template <typename DESCRIPTOR, typename BULK, bool FORCED>
static void runCase(const CaseSpec& c, const std::string& dir)
{
  // One cuboid of exactly n[0] x n[1] x n[2] cells, spacing 1, split into
  // several with splitRegular when the case asks for it.
  CuboidDecomposition3D<T> cuboids(Vector<T,3>{0, 0, 0}, T{1},
                                   Vector<int,3>{c.n[0], c.n[1], c.n[2]}, 1);
  if (c.split > 0) {
    cuboids.splitRegular(0, c.split);
  }
  cuboids.setPeriodicity({bool(c.periodic[0]), bool(c.periodic[1]), bool(c.periodic[2])});
  BlockLoadBalancer<T> load(0, 1, cuboids.size(), 0);
  // overlap 1 by default: the halo depth the time step actually refreshes
  // (superLattice.hh:138-143)
  SuperGeometry<T,3> geometry(cuboids, load, c.overlap > 0 ? c.overlap : 1);
  forEachCoreCell(geometry, [&](int iC, LatticeR<3> local, LatticeR<3> global) {
    geometry.getBlockGeometry(iC).set(local, int(c.ids[coreIndex(c, global[0], global[1], global[2])]));
  });
  // Materials were written into the block directly, which does not flag the
  // super geometry for communication; rename(0, 0) changes no material but
  // sets the flag (superGeometry.hh:323-333), so communicate() then fills the
  // overlap -- including its periodic images -- as a geometry built with
  // SuperGeometry::rename would have it.
  geometry.rename(0, 0);
  geometry.communicate();

  if (std::getenv("OLB_REF_NORMALS")) {
    if (cuboids.size() != 1) throw std::runtime_error("normals are dumped for one cuboid only");
    dumpNormals(c, geometry, dir);
  }

  SuperLattice<T,DESCRIPTOR> lattice(geometry);
  lattice.statisticsOff();
  lattice.template defineDynamics<BULK>(geometry, 1);
  if (!c.local.count(2)) {
    lattice.template defineDynamics<BounceBack<T,DESCRIPTOR>>(geometry, 2);
  }
  for (const auto& w : c.walls) {
    if (c.local.count(w.id)) continue;
    lattice.template defineDynamics<BounceBackVelocity<T,DESCRIPTOR>>(geometry, w.id);
  }
  for (const auto& [id, kind] : c.local) {
    // the boundaries' own default mixin, BGKdynamics, whatever the bulk
    // dynamics: boundary::set<boundary::InterpolatedVelocity>(lattice,
    // geometry, m) as aorta3d.cpp:118-127 writes it next to a Smagorinsky bulk
    using MIXIN = BGKdynamics<T,DESCRIPTOR>;
    if (kind == "velocity") {
      boundary::set<boundary::LocalVelocity<T,DESCRIPTOR,MIXIN>>(lattice, geometry, id);
    } else if (kind == "pressure") {
      boundary::set<boundary::LocalPressure<T,DESCRIPTOR,MIXIN>>(lattice, geometry, id);
    } else if (kind == "interp_velocity") {
      boundary::set<boundary::InterpolatedVelocity<T,DESCRIPTOR,MIXIN>>(lattice, geometry, id);
    } else {
      boundary::set<boundary::InterpolatedPressure<T,DESCRIPTOR,MIXIN>>(lattice, geometry, id);
    }
  }
  lattice.template setParameter<descriptors::OMEGA>(T(c.omega));
  if (c.dyn == "trt" || c.dyn == "forced_trt") {
    lattice.template setParameter<collision::TRT::MAGIC>(T(c.magic));
  }
  if (c.dyn == "smagorinsky") {
    lattice.template setParameter<collision::LES::SMAGORINSKY>(T(c.smago));
  }

  // Per-cell rho/u in the same order as the example programs: define the
  // moments, then initialise populations at equilibrium.
  const std::size_t nCore = c.ids.size();
  auto forEachInitCell = [&](auto f) {
    if (c.overlap > 0) forEachSpatialCell(c, geometry, c.overlap, f);
    else forEachCoreCell(geometry, f);
  };
  forEachInitCell([&](int iC, LatticeR<3> local, LatticeR<3> global) {
    const std::size_t k = coreIndex(c, global[0], global[1], global[2]);
    const int id = c.ids[k];
    if (id == 0) return;
    T rho = T(c.rhou[k]);
    T u[3] = {T(c.rhou[nCore + k]), T(c.rhou[2 * nCore + k]), T(c.rhou[3 * nCore + k])};
    auto cell = lattice.getBlock(iC).get(local);
    if constexpr (FORCED) {
      if (id == 1) {
        if (c.forceField.empty()) {
          cell.template setField<descriptors::FORCE>(
            Vector<T,3>{T(c.force[0]), T(c.force[1]), T(c.force[2])});
        } else {
          cell.template setField<descriptors::FORCE>(
            Vector<T,3>{T(c.forceField[k]), T(c.forceField[nCore + k]), T(c.forceField[2 * nCore + k])});
        }
      }
    }
    cell.defineRhoU(rho, u);
    cell.iniEquilibrium(rho, u);
  });
  // Moving walls: fixed density and wall velocity (cavity3d.cpp:124-136).
  for (const auto& w : c.walls) {
    forEachInitCell([&](int iC, LatticeR<3> local, LatticeR<3> global) {
      if (c.ids[coreIndex(c, global[0], global[1], global[2])] != w.id) return;
      auto cell = lattice.getBlock(iC).get(local);
      T u[3] = {T(w.u[0]), T(w.u[1]), T(w.u[2])};
      cell.defineRhoU(T(w.rho), u);
    });
  }
  lattice.initialize();

  if (std::getenv("OLB_REF_TIMING")) {
    // timing mode, as the benchmark harness does it (cavity3d.cpp:208-243):
    // 10 warm-up steps, then steps[0] timed steps, MLUPs over cells with dynamics
    for (int i = 0; i < 10; ++i) lattice.collideAndStream();
    std::size_t cells = 0;
    for (unsigned char id : c.ids) cells += id != 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (long i = 0; i < c.steps.at(0); ++i) lattice.collideAndStream();
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("{\"cells\": %zu, \"steps\": %ld, \"seconds\": %.6f, \"mlups\": %.3f}\n", cells, c.steps.at(0), s,
                double(cells) * double(c.steps.at(0)) / s / 1e6);
    return;
  }

  // Statistics mode: OpenLB accumulates the statistics of step k during its
  // collision, i.e. from the populations after k-1 steps; the first step's
  // averages are OpenLB's placeholders (latticeStatistics.hh:64-72).
  const bool stats = std::getenv("OLB_REF_STATS") != nullptr;
  if (stats) lattice.statisticsOn();
  const bool save = std::getenv("OLB_REF_SAVE") != nullptr;
  const bool vtk = std::getenv("OLB_REF_VTK") != nullptr;
  if (const char* restart = std::getenv("OLB_REF_LOAD")) {
    if (!lattice.template load<false>(restart)) throw std::runtime_error(std::string("cannot load ") + restart);
  }
  long done = 0;
  for (long target : c.steps) {
    for (; done < target; ++done) {
      lattice.collideAndStream();
    }
    dumpPopulations<DESCRIPTOR>(c, geometry, lattice, dir, target);
    if (save) {
      lattice.template save<false>(dir + "/checkpoint_" + std::to_string(target));
    }
    if (stats) {
      auto& st = lattice.getStatistics();
      const std::string name = dir + "/stats_" + std::to_string(target) + ".txt";
      std::FILE* f = std::fopen(name.c_str(), "w");
      if (!f) throw std::runtime_error("cannot write " + name);
      std::fprintf(f, "%zu %.17g %.17g %.17g\n", std::size_t(st.getNumCells()), double(st.getAverageRho()),
                   double(st.getAverageEnergy()), double(st.getMaxU()));
      std::fclose(f);
    }
  }
  if (vtk) {
    singleton::directories().setOutputDir(dir + "/");
    const UnitConverter<T,DESCRIPTOR> converter(T(0.01), T(1e-4), T(0.1), T(1), T(1e-5), T(1000), T(100));
    SuperVTMwriter3D<T> writer("fields");
    SuperLatticeDensity3D<T,DESCRIPTOR> density(lattice);
    SuperLatticeVelocity3D<T,DESCRIPTOR> velocity(lattice);
    SuperLatticePhysVelocity3D<T,DESCRIPTOR> physVelocity(lattice, converter);
    SuperLatticePhysPressure3D<T,DESCRIPTOR> physPressure(lattice, converter);
    writer.addFunctor(density);
    writer.addFunctor(velocity);
    writer.addFunctor(physVelocity);
    writer.addFunctor(physPressure);
    writer.createMasterFile();
    writer.write(int(done));
  }
}

int main(int argc, char** argv)
{
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <case-dir>\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  initialize(&argc, &argv, false, false);
  const CaseSpec c = readCase(dir);
  using D19 = descriptors::D3Q19<>;
  using D19F = descriptors::D3Q19<descriptors::FORCE>;
  using D27 = descriptors::D3Q27<>;
  using D27F = descriptors::D3Q27<descriptors::FORCE>;
  if (c.q == 19) {
    if (c.dyn == "bgk") runCase<D19, BGKdynamics<T,D19>, false>(c, dir);
    else if (c.dyn == "forced_bgk") runCase<D19F, ForcedBGKdynamics<T,D19F>, true>(c, dir);
    else if (c.dyn == "trt") runCase<D19, TRTdynamics<T,D19>, false>(c, dir);
    else if (c.dyn == "forced_trt") runCase<D19F, ForcedTRTdynamics<T,D19F>, true>(c, dir);
    else if (c.dyn == "smagorinsky") runCase<D19, SmagorinskyBGKdynamics<T,D19>, false>(c, dir);
    else throw std::runtime_error("unknown dyn " + c.dyn);
  } else if (c.q == 27) {
    if (c.dyn == "bgk") runCase<D27, BGKdynamics<T,D27>, false>(c, dir);
    else if (c.dyn == "forced_bgk") runCase<D27F, ForcedBGKdynamics<T,D27F>, true>(c, dir);
    else if (c.dyn == "trt") runCase<D27, TRTdynamics<T,D27>, false>(c, dir);
    else if (c.dyn == "forced_trt") runCase<D27F, ForcedTRTdynamics<T,D27F>, true>(c, dir);
    else if (c.dyn == "smagorinsky") runCase<D27, SmagorinskyBGKdynamics<T,D27>, false>(c, dir);
    else throw std::runtime_error("unknown dyn " + c.dyn);
  } else {
    throw std::runtime_error("q must be 19 or 27");
  }
  return 0;
}
