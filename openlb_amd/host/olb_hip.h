// olb_hip.h -- C++ host layer of the MI355X block-lattice backend.
//
// Header-only.  Mirrors the part of OpenLB's API that drives the time step
// (names, argument meaning, error behaviour), on top of the C ABI in lbmh.h:
//
//   SuperLattice<T,DESCRIPTOR>(extent, periodicity)      superLattice.h:84-101
//   defineDynamics<DYNAMICS>(indicator)                  superLattice.h:214-233
//   setParameter<PARAMETER>(value)                       superLattice.h:318-336
//   defineRhoU / iniEquilibrium(indicator, rho, u)       superLattice.h:344-352
//   collideAndStream()                                   superLattice.h:380
//   setProcessingContext(Evaluation | Simulation)        superLattice.h:154-164
//   get(x, y, z)[iPop] on the host mirror                superLattice.h:169-176
//   statisticsOff()                                      superLattice.h:397-403
//   boundary::set<Local|InterpolatedVelocity/Pressure>   boundary/setBoundary.h:83-101
//   getStatistics()                                      superLattice.h:378
//   getSerializableSize / save / load (buffer or file)   core/serializer.h:145-189
//   Ensemble: many set-ups advanced in one launch per step (no OpenLB counterpart)
//   UnitConverter, SuperLatticeDensity3D / Velocity3D /
//   PhysVelocity3D / PhysPressure3D, SuperVTMwriter3D    core/unitConverter.h, functors/lattice/, io/superVtmWriter3D.h
//
// Dynamics are tags carrying the C ABI's bulk kind; parameters are tags of
// the same names as OpenLB's fields.  HIP errors surface as std::runtime_error
// (like gpu::cuda::device::check, src/core/platform/gpu/cuda/device.hh:48-62),
// invalid requests as std::invalid_argument / std::logic_error.
#pragma once

#include <zlib.h>

#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iterator>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "lbmh.h"

namespace olb {
namespace hip {

// ---- descriptors (src/descriptor/definition/common.h:120-166, 313-366) ----
namespace descriptors {
struct D3Q19 { static constexpr int d = 3, q = 19; };
struct D3Q27 { static constexpr int d = 3, q = 27; };
// parameter tags (descriptors::OMEGA, descriptors::FORCE)
struct OMEGA {};
struct FORCE {};
}  // namespace descriptors

namespace collision {
struct TRT { struct MAGIC {}; };
namespace LES { struct SMAGORINSKY {}; }
}  // namespace collision

// ---- dynamics tags (src/dynamics/dynamics.h, smagorinskyBGKdynamics.h) ----
enum class DynamicsRole { Bulk, BounceBack, BounceBackVelocity, None };

template <typename T, typename D> struct BGKdynamics { static constexpr int kind = LBMH_BGK; static constexpr DynamicsRole role = DynamicsRole::Bulk; };
template <typename T, typename D> struct ForcedBGKdynamics { static constexpr int kind = LBMH_FORCED_BGK; static constexpr DynamicsRole role = DynamicsRole::Bulk; };
template <typename T, typename D> struct TRTdynamics { static constexpr int kind = LBMH_TRT; static constexpr DynamicsRole role = DynamicsRole::Bulk; };
template <typename T, typename D> struct ForcedTRTdynamics { static constexpr int kind = LBMH_FORCED_TRT; static constexpr DynamicsRole role = DynamicsRole::Bulk; };
template <typename T, typename D> struct SmagorinskyBGKdynamics { static constexpr int kind = LBMH_SMAGORINSKY; static constexpr DynamicsRole role = DynamicsRole::Bulk; };
template <typename T, typename D> struct BounceBack { static constexpr DynamicsRole role = DynamicsRole::BounceBack; };
template <typename T, typename D> struct BounceBackVelocity { static constexpr DynamicsRole role = DynamicsRole::BounceBackVelocity; };
template <typename T, typename D> struct NoDynamics { static constexpr DynamicsRole role = DynamicsRole::None; };

// ---- local boundaries (src/boundary/localVelocity3D.h, localPressure3D.h):
// regularized velocity / pressure boundaries with the bulk dynamics mixed in
namespace boundary {
template <typename T, typename D, typename MIXIN> struct LocalVelocity { static constexpr int kind = LBMH_LOCAL_VELOCITY; };
template <typename T, typename D, typename MIXIN> struct LocalPressure { static constexpr int kind = LBMH_LOCAL_PRESSURE; };
template <typename T, typename D, typename MIXIN> struct InterpolatedVelocity { static constexpr int kind = LBMH_INTERP_VELOCITY; };
template <typename T, typename D, typename MIXIN> struct InterpolatedPressure { static constexpr int kind = LBMH_INTERP_PRESSURE; };
}  // namespace boundary

enum class ProcessingContext { Evaluation, Simulation };

using Indicator = std::function<bool(int, int, int)>;
using ScalarField = std::function<double(int, int, int)>;
using VectorField = std::function<std::array<double, 3>(int, int, int)>;

namespace detail {
inline void check(int rc)
{
  if (rc == LBMH_OK) return;
  const std::string msg = lbmh_last_error();
  if (rc == LBMH_E_INVALID) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

template <typename T> constexpr int storage_of()
{
  if constexpr (std::is_same_v<T, double>) return LBMH_F64;
  else if constexpr (std::is_same_v<T, float>) return LBMH_F32;
  else return LBMH_F16;
}

// base64 with the standard alphabet and '=' padding, as OpenLB's
// Base64Encoder / Base64Decoder write and read it (src/io/base64.hh:45-194)
inline std::string base64_encode(const std::uint8_t* p, std::size_t n)
{
  static const char enc[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  for (std::size_t i = 0; i < n; i += 3) {
    const std::uint32_t b0 = p[i], b1 = i + 1 < n ? p[i + 1] : 0u, b2 = i + 2 < n ? p[i + 2] : 0u;
    const std::uint32_t v = (b0 << 16) | (b1 << 8) | b2;
    out += enc[(v >> 18) & 63];
    out += enc[(v >> 12) & 63];
    out += i + 1 < n ? enc[(v >> 6) & 63] : '=';
    out += i + 2 < n ? enc[v & 63] : '=';
  }
  return out;
}

// Decodes `bytes` bytes from the base64 text at `pos`, skipping white space;
// advances `pos` past the groups it used.
inline bool base64_decode(const std::string& text, std::size_t& pos, std::uint8_t* out, std::size_t bytes)
{
  auto value = [](char c) -> int {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+') return 62;
    if (c == '/') return 63;
    if (c == '=') return 0;
    return -1;
  };
  std::size_t done = 0;
  while (done < bytes) {
    int v[4];
    for (int k = 0; k < 4; ++k) {
      while (pos < text.size() && std::isspace(static_cast<unsigned char>(text[pos]))) ++pos;
      if (pos >= text.size() || (v[k] = value(text[pos])) < 0) return false;
      ++pos;
    }
    const std::uint32_t w = (std::uint32_t(v[0]) << 18) | (std::uint32_t(v[1]) << 12) | (std::uint32_t(v[2]) << 6) | v[3];
    const std::uint8_t b[3] = {std::uint8_t(w >> 16), std::uint8_t(w >> 8), std::uint8_t(w)};
    for (int k = 0; k < 3 && done < bytes; ++k) out[done++] = b[k];
  }
  return true;
}

// createParallelFileName(name) + ".dat" of a one-rank run (src/io/fileName.hh:51-64)
inline std::string parallel_file_name(const std::string& name) { return name + "_rank0_size1.dat"; }

template <typename DYN> struct kind_of { static constexpr int value = -1; };
template <template <typename, typename> class DYN, typename T, typename D>
struct kind_of<DYN<T, D>> {
  static constexpr int value = []() {
    if constexpr (DYN<T, D>::role == DynamicsRole::Bulk) return DYN<T, D>::kind;
    else return -1;
  }();
};
}  // namespace detail

// Host view of one cell of the evaluation mirror (cell[iPop] as in cpu::Cell).
template <typename T>
class ConstCellView {
 public:
  ConstCellView(const std::vector<double>& pop, std::size_t n, std::size_t c) : _pop(pop), _n(n), _c(c) {}
  T operator[](unsigned iPop) const { return T(_pop[iPop * _n + _c]); }

 private:
  const std::vector<double>& _pop;
  std::size_t _n, _c;
};

// One block of OpenLB's super lattice, resident on one MI355X.  Storage type
// follows T (double / float); FP16 storage with FP32 arithmetic is selected
// with T = float and storage = LBMH_F16.
template <typename T, typename DESCRIPTOR>
class Ensemble;

template <typename T, typename DESCRIPTOR>
class SuperLattice {
  friend class Ensemble<T, DESCRIPTOR>;

 public:
  static constexpr int q = DESCRIPTOR::q;

  SuperLattice(std::array<int, 3> extent, std::array<bool, 3> periodicity, int device = 0,
               int storage = detail::storage_of<T>())
      : _extent(extent), _periodic(periodicity), _device(device), _storage(storage),
        _ids(std::size_t(extent[0]) * extent[1] * extent[2], LBMH_ID_NONE)
  {
    if (storage == LBMH_F16 && !std::is_same_v<T, float>) {
      throw std::invalid_argument("FP16 storage computes in FP32: use T = float");
    }
  }

  SuperLattice(const SuperLattice&) = delete;
  SuperLattice& operator=(const SuperLattice&) = delete;
  ~SuperLattice() { lbmh_lattice_destroy(_handle); }

  std::size_t ncore() const { return _ids.size(); }
  const std::array<int, 3>& extent() const { return _extent; }

  /// Set dynamics of the indicated cells (SuperLattice::defineDynamics).
  template <typename DYNAMICS>
  void defineDynamics(const Indicator& indicator)
  {
    requireSetup("defineDynamics");
    uint8_t id = LBMH_ID_NONE;
    if constexpr (DYNAMICS::role == DynamicsRole::Bulk) {
      if (_kind >= 0 && _kind != DYNAMICS::kind) {
        throw std::logic_error("one bulk dynamics per block lattice");
      }
      _kind = DYNAMICS::kind;
      id = LBMH_ID_BULK;
    } else if constexpr (DYNAMICS::role == DynamicsRole::BounceBack) {
      id = LBMH_ID_BOUNCE_BACK;
    } else if constexpr (DYNAMICS::role == DynamicsRole::BounceBackVelocity) {
      id = uint8_t(LBMH_ID_FIRST_WALL + _wallsDefined);
      if (id == 0) throw std::logic_error("at most 253 moving-wall groups per block");
      _pendingWall = id;
      ++_wallsDefined;
    }
    forEach([&](int x, int y, int z, std::size_t c) {
      if (indicator(x, y, z)) _ids[c] = id;
    });
  }

  /// Update PARAMETER of the bulk dynamics (SuperLattice::setParameter).
  template <typename PARAMETER>
  void setParameter(T value)
  {
    if constexpr (std::is_same_v<PARAMETER, descriptors::OMEGA>) _params.omega = double(value);
    else if constexpr (std::is_same_v<PARAMETER, collision::TRT::MAGIC>) _params.magic = double(value);
    else if constexpr (std::is_same_v<PARAMETER, collision::LES::SMAGORINSKY>) _params.smago = double(value);
    else static_assert(!sizeof(PARAMETER), "unsupported parameter");
    if (_handle) detail::check(lbmh_set_bulk_params(_handle, &_params));
  }

  template <typename PARAMETER>
  void setParameter(std::array<T, 3> value)
  {
    static_assert(std::is_same_v<PARAMETER, descriptors::FORCE>, "vector parameter must be FORCE");
    for (int d = 0; d < 3; ++d) _params.force[d] = double(value[d]);
    if (_handle) detail::check(lbmh_set_bulk_params(_handle, &_params));
  }

  /// Local boundary BC on the indicated cells (boundary::set, setBoundary.h:
  /// 83-101); BC is boundary::LocalVelocity or boundary::LocalPressure with
  /// the bulk dynamics as mixin.  The cells' types and normals are found from
  /// the dynamics defined so far (bulk = fluid, no dynamics = outside) when the
  /// lattice is initialized.  defineWall() then gives the velocity (density).
  template <typename BC>
  void setLocalBoundary(const Indicator& indicator)
  {
    requireSetup("boundary::set");
    const uint8_t id = uint8_t(LBMH_ID_FIRST_WALL + _wallsDefined);
    if (id == 0) throw std::logic_error("at most 253 wall and boundary groups per block");
    ++_wallsDefined;
    _pendingWall = id;
    _local.push_back({int(id), BC::kind});
    forEach([&](int x, int y, int z, std::size_t c) {
      if (indicator(x, y, z)) _ids[c] = id;
    });
  }

  /// Fixed density and velocity of the wall or boundary group defined last
  /// (defineRhoU / defineU on its material, cavity3d.cpp:113-131).
  void defineWall(T rho, std::array<T, 3> u)
  {
    if (!_pendingWall) throw std::logic_error("defineWall needs a wall or boundary group first");
    _walls.push_back({int(_pendingWall), double(rho), {double(u[0]), double(u[1]), double(u[2])}});
    if (_handle) applyWalls();
  }

  /// Equilibrium at (rho, u) on the indicated cells (SuperLattice::iniEquilibrium).
  void iniEquilibrium(const Indicator& indicator, const ScalarField& rho, const VectorField& u)
  {
    const std::size_t n = ncore();
    if (_rhoU.empty()) {
      _rhoU.assign(4 * n, 0.0);
      for (std::size_t c = 0; c < n; ++c) _rhoU[c] = 1.0;
    }
    forEach([&](int x, int y, int z, std::size_t c) {
      if (!indicator(x, y, z)) return;
      _rhoU[c] = rho(x, y, z);
      const auto v = u(x, y, z);
      for (int d = 0; d < 3; ++d) _rhoU[(d + 1) * n + c] = v[d];
    });
    if (_handle) detail::check(lbmh_ini_equilibrium(_handle, _rhoU.data()));
  }

  /// Allocate the device block and upload ids, parameters and the initial
  /// state (SuperLattice::initialize + setProcessingContext(Simulation)).
  void initialize()
  {
    if (_handle) return;
    lbmh_lattice_desc desc{};
    desc.q = q;
    desc.nx = _extent[0];
    desc.ny = _extent[1];
    desc.nz = _extent[2];
    for (int d = 0; d < 3; ++d) desc.periodic[d] = _periodic[d] ? 1 : 0;
    desc.storage = _storage;
    desc.bulk_kind = _kind >= 0 ? _kind : LBMH_BGK;
    desc.device = _device;
    detail::check(lbmh_lattice_create(&desc, &_handle));
    detail::check(lbmh_set_dynamics_ids(_handle, _ids.data()));
    detail::check(lbmh_set_bulk_params(_handle, &_params));
    applyWalls();
    for (const auto& b : _local) {
      Wall w{b.id, 1.0, {0.0, 0.0, 0.0}};
      for (const auto& v : _walls)
        if (v.id == b.id) w = v;
      detail::check(lbmh_set_local_boundary(_handle, b.id, b.kind, w.rho, w.u));
    }
    if (!_rhoU.empty()) detail::check(lbmh_ini_equilibrium(_handle, _rhoU.data()));
    // initialize(): the post-processors of the local boundaries run once
    detail::check(lbmh_initialize(_handle));
  }

  /// One time step (SuperLattice::collideAndStream), asynchronous.
  void collideAndStream(void* stream = nullptr)
  {
    if (!_handle) initialize();
    detail::check(lbmh_step(_handle, 1, stream));
    _mirrorValid = false;
  }

  /// Evaluation: download populations to the host mirror; Simulation: upload it.
  void setProcessingContext(ProcessingContext context)
  {
    if (!_handle) initialize();
    _mirror.resize(std::size_t(q) * ncore());
    if (context == ProcessingContext::Evaluation) {
      detail::check(lbmh_populations_download(_handle, _mirror.data(), LBMH_F64));
      _mirrorValid = true;
    } else {
      if (!_mirrorValid) throw std::logic_error("no host populations to upload");
      detail::check(lbmh_populations_upload(_handle, _mirror.data(), LBMH_F64));
    }
  }

  ConstCellView<T> get(int x, int y, int z) const
  {
    if (!_mirrorValid) throw std::logic_error("call setProcessingContext(Evaluation) first");
    return ConstCellView<T>(_mirror, ncore(), index(x, y, z));
  }

  /// Mutable host mirror, [q][nx][ny][nz], shifted values.
  std::vector<double>& populations() { return _mirror; }

  void synchronize() const { if (_handle) detail::check(lbmh_synchronize(_handle)); }
  void statisticsOff() {}
  void statisticsOn() {}

  /// Lattice statistics, computed on demand on the device in FP64 (no cost in
  /// the time step).  OpenLB's getStatistics() after step k reports the
  /// populations the collision of step k saw; this reports the current ones,
  /// i.e. what step k+1 will see (src/core/latticeStatistics.h:49-71).
  struct Statistics {
    std::size_t numCells;
    double averageRho, averageEnergy, maxU;
    double getAverageRho() const { return averageRho; }
    double getAverageEnergy() const { return averageEnergy; }
    double getMaxU() const { return maxU; }
    std::size_t getNumCells() const { return numCells; }
  };

  Statistics getStatistics() const
  {
    double out[4];
    detail::check(lbmh_statistics(_handle, out));
    return Statistics{std::size_t(out[0]), out[1], out[2], out[3]};
  }

  /// FP64 {sum rho, sum j} over bulk cells.
  std::array<double, 4> conservedMoments() const
  {
    std::array<double, 4> out{};
    detail::check(lbmh_reduce_moments(_handle, out.data()));
    return out;
  }

  std::size_t index(int x, int y, int z) const
  {
    return (std::size_t(x) * _extent[1] + y) * _extent[2] + z;
  }

  lbmh_lattice* handle() const { return _handle; }

  /// Serializable interface (src/core/serializer.h:145-189) in OpenLB's value
  /// type T: the bytes Serializable::save(std::uint8_t*) writes for an OpenLB
  /// block of the same set-up (FP16 storage converts to float).  A load checks
  /// the field values the dynamics read against this lattice's set-up and
  /// throws std::invalid_argument on a mismatch.
  std::size_t getSerializableSize()
  {
    if (!_handle) initialize();
    std::size_t n = 0;
    detail::check(lbmh_checkpoint_size(_handle, real(), &n));
    return n;
  }

  bool save(std::uint8_t* buffer)
  {
    const std::size_t n = getSerializableSize();
    detail::check(lbmh_checkpoint_save(_handle, real(), buffer, n));
    return true;
  }

  bool load(const std::uint8_t* buffer)
  {
    const std::size_t n = getSerializableSize();
    detail::check(lbmh_checkpoint_load(_handle, real(), buffer, n));
    _mirrorValid = false;
    return true;
  }

  /// The file form: <logOutDir><fileName>_rank0_size1.dat holding the base64
  /// of the byte count (size_t, or unsigned with enforceUint) and of the
  /// bytes (src/core/serializer.hh:81-97, io/serializerIO.hh:40-94).  The
  /// log output directory is OpenLB's default, "./tmp/", unless set with
  /// setLogOutDir; includeLogOutputDir = false writes next to the program.
  template <bool includeLogOutputDir = true>
  bool save(std::string fileName = "", bool enforceUint = false)
  {
    std::vector<std::uint8_t> raw(getSerializableSize());
    save(raw.data());
    std::string text;
    if (enforceUint) {
      if (raw.size() > 0xffffffffu) return false;
      const std::uint32_t size = std::uint32_t(raw.size());
      text = detail::base64_encode(reinterpret_cast<const std::uint8_t*>(&size), sizeof(size));
    } else {
      const std::uint64_t size = raw.size();
      text = detail::base64_encode(reinterpret_cast<const std::uint8_t*>(&size), sizeof(size));
    }
    text += detail::base64_encode(raw.data(), raw.size());
    std::ofstream out(fullFileName<includeLogOutputDir>(fileName));
    if (!out) return false;
    out << text;
    return bool(out);
  }

  template <bool includeLogOutputDir = true>
  bool load(std::string fileName = "", bool enforceUint = false)
  {
    std::ifstream in(fullFileName<includeLogOutputDir>(fileName));
    if (!in) return false;
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::size_t pos = 0;
    std::uint64_t size = 0;
    if (enforceUint) {
      std::uint32_t s32 = 0;
      if (!detail::base64_decode(text, pos, reinterpret_cast<std::uint8_t*>(&s32), sizeof(s32))) return false;
      size = s32;
    } else if (!detail::base64_decode(text, pos, reinterpret_cast<std::uint8_t*>(&size), sizeof(size))) {
      return false;
    }
    if (size != getSerializableSize()) {
      throw std::invalid_argument("checkpoint " + fileName + " holds " + std::to_string(size) +
                                  " bytes, this lattice " + std::to_string(getSerializableSize()));
    }
    std::vector<std::uint8_t> raw(size);
    if (!detail::base64_decode(text, pos, raw.data(), raw.size())) return false;
    return load(raw.data());
  }

  static void setLogOutDir(std::string dir) { logOutDir() = std::move(dir); }

 private:
  static std::string& logOutDir()
  {
    static std::string dir = "./tmp/";
    return dir;
  }

  template <bool includeLogOutputDir>
  static std::string fullFileName(std::string fileName)
  {
    if (fileName.empty()) fileName = "Serializable";
    return (includeLogOutputDir ? logOutDir() : std::string()) + detail::parallel_file_name(fileName);
  }

  int real() const { return _storage == LBMH_F64 ? LBMH_F64 : LBMH_F32; }

  struct Wall { int id; double rho; double u[3]; };
  struct Local { int id; int kind; };

  template <typename F>
  void forEach(F&& f) const
  {
    for (int x = 0; x < _extent[0]; ++x)
      for (int y = 0; y < _extent[1]; ++y)
        for (int z = 0; z < _extent[2]; ++z) f(x, y, z, index(x, y, z));
  }

  void requireSetup(const char* what) const
  {
    if (_handle) throw std::logic_error(std::string(what) + " after initialize()");
  }

  bool isLocal(int id) const
  {
    for (const auto& b : _local)
      if (b.id == id) return true;
    return false;
  }

  void applyWalls()
  {
    for (const auto& w : _walls)
      if (!isLocal(w.id)) detail::check(lbmh_set_wall(_handle, w.id, w.rho, w.u));
  }

  std::array<int, 3> _extent;
  std::array<bool, 3> _periodic;
  int _device;
  int _storage;
  int _kind = -1;
  std::vector<uint8_t> _ids;
  std::vector<double> _rhoU;
  std::vector<double> _mirror;
  bool _mirrorValid = false;
  lbmh_bulk_params _params{1.0, 0.25, 0.0, {0.0, 0.0, 0.0}};
  std::vector<Wall> _walls;
  std::vector<Local> _local;
  uint8_t _pendingWall = 0;
  int _wallsDefined = 0;
  lbmh_lattice* _handle = nullptr;
};

namespace boundary {
/// boundary::set<BC>(lattice, indicator)           src/boundary/setBoundary.h:83-101
template <typename BC, typename LATTICE>
void set(LATTICE& lattice, const Indicator& indicator)
{
  lattice.template setLocalBoundary<BC>(indicator);
}
}  // namespace boundary

/// Many independent lattices advanced together, one launch per step for all
/// of them (lbmh_ensemble_*).  OpenLB runs such sets (UQ samples, parameter
/// sweeps) as one SuperLattice after the other; here every member is set up
/// with the SuperLattice calls above -- defineDynamics, setParameter,
/// boundary::set, defineWall, iniEquilibrium -- and handed to add() instead of
/// initialize(), so no device lattice is created for it.  Members of one
/// ensemble share the descriptor, the storage type and the bulk dynamics; each
/// has its own extent, parameters, walls (ids < 64) and step budget.
template <typename T, typename DESCRIPTOR>
class Ensemble {
 public:
  Ensemble(int maxMembers, std::size_t populationBytes, std::size_t cellCapacity, int bulkKind = LBMH_BGK,
           int device = 0, int storage = detail::storage_of<T>())
  {
    lbmh_ensemble_desc d{};
    d.q = DESCRIPTOR::q;
    d.storage = storage;
    d.bulk_kind = bulkKind;
    d.device = device;
    d.max_members = maxMembers;
    d.population_bytes = populationBytes;
    d.cell_capacity = cellCapacity;
    d.boundary_capacity = 0;
    detail::check(lbmh_ensemble_create(&d, &_handle));
    _extents.resize(std::size_t(maxMembers));
  }
  Ensemble(const Ensemble&) = delete;
  Ensemble& operator=(const Ensemble&) = delete;
  ~Ensemble() { lbmh_ensemble_destroy(_handle); }

  /// Device bytes a member of this extent takes (populations, ids): the admission check.
  std::pair<std::size_t, std::size_t> memberBytes(std::array<int, 3> extent) const
  {
    std::size_t pop = 0, ids = 0;
    detail::check(lbmh_ensemble_member_bytes(_handle, extent.data(), &pop, &ids));
    return {pop, ids};
  }

  /// Admit the set-up of `lattice` as a member that runs `steps` steps (< 0:
  /// until removed); returns the member id.  The lattice itself stays
  /// uninitialized and may be reused for the next member.
  int add(const SuperLattice<T, DESCRIPTOR>& lattice, long steps = -1)
  {
    if (lattice._handle) throw std::logic_error("add() takes a lattice set-up that was not initialized");
    int periodic[3];
    for (int d = 0; d < 3; ++d) periodic[d] = lattice._periodic[d] ? 1 : 0;
    int m = -1;
    detail::check(lbmh_ensemble_add(_handle, lattice._extent.data(), periodic, lattice._ids.data(), &lattice._params,
                                    int64_t(steps), &m));
    for (const auto& w : lattice._walls)
      if (!lattice.isLocal(w.id)) detail::check(lbmh_ensemble_set_wall(_handle, m, w.id, w.rho, w.u));
    for (const auto& b : lattice._local) {
      double rho = 1.0, u[3] = {0.0, 0.0, 0.0};
      for (const auto& w : lattice._walls)
        if (w.id == b.id) {
          rho = w.rho;
          for (int d = 0; d < 3; ++d) u[d] = w.u[d];
        }
      detail::check(lbmh_ensemble_set_local_boundary(_handle, m, b.id, b.kind, rho, u));
    }
    if (!lattice._rhoU.empty()) detail::check(lbmh_ensemble_ini_equilibrium(_handle, m, lattice._rhoU.data()));
    if (!lattice._local.empty()) detail::check(lbmh_ensemble_initialize(_handle, m));
    _extents[std::size_t(m)] = lattice._extent;
    return m;
  }

  void remove(int member) { detail::check(lbmh_ensemble_remove(_handle, member)); }
  void setSteps(int member, long steps) { detail::check(lbmh_ensemble_set_steps(_handle, member, int64_t(steps))); }

  /// {steps done, steps left} of a member (synchronises the device).
  std::pair<long, long> status(int member)
  {
    int64_t done = 0, left = 0;
    detail::check(lbmh_ensemble_status(_handle, member, &done, &left));
    return {long(done), long(left)};
  }

  /// n steps of every member that has steps left, asynchronously; with graph,
  /// the step loop is replayed from a captured HIP graph.
  void collideAndStream(long n = 1, bool graph = true, void* stream = nullptr)
  {
    detail::check(graph ? lbmh_ensemble_step_graph(_handle, int64_t(n), stream)
                        : lbmh_ensemble_step(_handle, int64_t(n), stream));
  }

  /// FP64 rho and u of a member, [4][nx][ny][nz].
  std::vector<double> moments(int member)
  {
    const auto& n = _extents.at(std::size_t(member));
    std::vector<double> out(4 * std::size_t(n[0]) * n[1] * n[2]);
    detail::check(lbmh_ensemble_moments(_handle, member, out.data()));
    return out;
  }

  /// Shifted populations of a member in FP64, [q][nx][ny][nz].
  std::vector<double> populations(int member)
  {
    const auto& n = _extents.at(std::size_t(member));
    std::vector<double> out(std::size_t(DESCRIPTOR::q) * n[0] * n[1] * n[2]);
    detail::check(lbmh_ensemble_download(_handle, member, out.data(), LBMH_F64));
    return out;
  }

  /// One member's OpenLB checkpoint bytes (Serializable::save / load of its block).
  std::vector<std::uint8_t> save(int member)
  {
    std::size_t n = 0;
    detail::check(lbmh_ensemble_checkpoint_size(_handle, member, real(), &n));
    std::vector<std::uint8_t> raw(n);
    detail::check(lbmh_ensemble_checkpoint_save(_handle, member, real(), raw.data(), raw.size()));
    return raw;
  }
  void load(int member, const std::vector<std::uint8_t>& raw)
  {
    detail::check(lbmh_ensemble_checkpoint_load(_handle, member, real(), raw.data(), raw.size()));
  }

  const std::array<int, 3>& extent(int member) const { return _extents.at(std::size_t(member)); }
  lbmh_ensemble* handle() const { return _handle; }

 private:
  int real() const { return std::is_same_v<T, double> ? LBMH_F64 : LBMH_F32; }

  lbmh_ensemble* _handle = nullptr;
  std::vector<std::array<int, 3>> _extents;
};

/// UnitConverter(physDeltaX, physDeltaT, charPhysLength, charPhysVelocity,
/// physViscosity, physDensity, charPhysPressure)       src/core/unitConverter.h:152-173
template <typename T, typename DESCRIPTOR>
class UnitConverter {
 public:
  UnitConverter(T physDeltaX, T physDeltaT, T charPhysLength, T charPhysVelocity, T physViscosity, T physDensity,
                T charPhysPressure = 0)
      : _conversionLength(physDeltaX), _conversionTime(physDeltaT),
        _conversionVelocity(_conversionLength / _conversionTime), _conversionDensity(physDensity),
        _conversionMass(_conversionDensity * std::pow(_conversionLength, 3)),
        _conversionViscosity(_conversionLength * _conversionLength / _conversionTime),
        _conversionForce(_conversionMass * _conversionLength / (_conversionTime * _conversionTime)),
        _conversionPressure(_conversionForce / std::pow(_conversionLength, 2)), _charPhysLength(charPhysLength),
        _charPhysVelocity(charPhysVelocity), _physViscosity(physViscosity), _charPhysPressure(charPhysPressure),
        _latticeRelaxationTime(_physViscosity / _conversionViscosity * T(3) + T(0.5))
  {
  }
  T getConversionFactorLength() const { return _conversionLength; }
  T getPhysDeltaX() const { return _conversionLength; }
  T getConversionFactorVelocity() const { return _conversionVelocity; }
  T getConversionFactorPressure() const { return _conversionPressure; }
  T getCharPhysPressure() const { return _charPhysPressure; }
  T getCharLatticeVelocity() const { return _charPhysVelocity / _conversionVelocity; }
  T getLatticeRelaxationTime() const { return _latticeRelaxationTime; }
  T getLatticeRelaxationFrequency() const { return T(1) / _latticeRelaxationTime; }
  T getPhysVelocity(T latticeVelocity) const { return _conversionVelocity * latticeVelocity; }
  T getPhysPressure(T latticePressure) const { return _conversionPressure * latticePressure + _charPhysPressure; }

 private:
  T _conversionLength, _conversionTime, _conversionVelocity, _conversionDensity, _conversionMass;
  T _conversionViscosity, _conversionForce, _conversionPressure, _charPhysLength, _charPhysVelocity;
  T _physViscosity, _charPhysPressure, _latticeRelaxationTime;
};

namespace detail {
/// A lattice functor the VTK writer evaluates on the device: on a lattice
/// (lbmh_field_export) or on one ensemble member (lbmh_ensemble_field_export).
template <typename T, typename DESCRIPTOR>
class LatticeFieldF {
 public:
  LatticeFieldF(SuperLattice<T, DESCRIPTOR>& lattice, int field, std::string name, const double* conversion)
      : _field(field), _name(std::move(name))
  {
    setConversion(conversion);
    SuperLattice<T, DESCRIPTOR>* l = &lattice;
    _extent = l->extent();
    _export = [l](int f, int overlap, const double* conv, float* out, std::size_t n) {
      if (!l->handle()) l->initialize();
      check(lbmh_field_export(l->handle(), f, overlap, conv, out, n));
    };
  }
  LatticeFieldF(Ensemble<T, DESCRIPTOR>& ensemble, int member, int field, std::string name, const double* conversion)
      : _field(field), _name(std::move(name))
  {
    setConversion(conversion);
    Ensemble<T, DESCRIPTOR>* e = &ensemble;
    _extent = e->extent(member);
    _export = [e, member](int f, int overlap, const double* conv, float* out, std::size_t n) {
      check(lbmh_ensemble_field_export(e->handle(), member, f, overlap, conv, out, n));
    };
  }
  virtual ~LatticeFieldF() = default;
  std::string& getName() { return _name; }
  int getTargetDim() const { return (_field == LBMH_FIELD_VELOCITY || _field == LBMH_FIELD_PHYS_VELOCITY) ? 3 : 1; }
  const std::array<int, 3>& extent() const { return _extent; }

  /// The writer's float buffer over [-overlap, extent + overlap), x fastest.
  std::vector<float> evaluate(int overlap)
  {
    const auto& n = _extent;
    const std::size_t points = std::size_t(n[0] + 2 * overlap) * (n[1] + 2 * overlap) * (n[2] + 2 * overlap);
    std::vector<float> out(points * std::size_t(getTargetDim()));
    _export(_field, overlap, _hasConversion ? _conversion : nullptr, out.data(), out.size());
    return out;
  }

 private:
  void setConversion(const double* conversion)
  {
    if (conversion)
      for (int k = 0; k < 3; ++k) _conversion[k] = conversion[k];
    _hasConversion = conversion != nullptr;
  }

  int _field;
  std::string _name;
  std::array<int, 3> _extent{};
  std::function<void(int, int, const double*, float*, std::size_t)> _export;
  double _conversion[3] = {0, 0, 0};
  bool _hasConversion = false;
};

template <typename T, typename DESCRIPTOR>
std::array<double, 3> conversion_of(const UnitConverter<T, DESCRIPTOR>& c)
{
  return {double(c.getConversionFactorVelocity()), double(c.getConversionFactorPressure()),
          double(c.getCharPhysPressure())};
}
}  // namespace detail

/// SuperLatticeDensity3D (src/functors/lattice/latticeDensity3D.hh)
template <typename T, typename DESCRIPTOR>
class SuperLatticeDensity3D : public detail::LatticeFieldF<T, DESCRIPTOR> {
 public:
  explicit SuperLatticeDensity3D(SuperLattice<T, DESCRIPTOR>& l)
      : detail::LatticeFieldF<T, DESCRIPTOR>(l, LBMH_FIELD_DENSITY, "density", nullptr) {}
  SuperLatticeDensity3D(Ensemble<T, DESCRIPTOR>& e, int member)
      : detail::LatticeFieldF<T, DESCRIPTOR>(e, member, LBMH_FIELD_DENSITY, "density", nullptr) {}
};

/// SuperLatticeVelocity3D (latticeVelocity3D.hh)
template <typename T, typename DESCRIPTOR>
class SuperLatticeVelocity3D : public detail::LatticeFieldF<T, DESCRIPTOR> {
 public:
  explicit SuperLatticeVelocity3D(SuperLattice<T, DESCRIPTOR>& l)
      : detail::LatticeFieldF<T, DESCRIPTOR>(l, LBMH_FIELD_VELOCITY, "velocity", nullptr) {}
  SuperLatticeVelocity3D(Ensemble<T, DESCRIPTOR>& e, int member)
      : detail::LatticeFieldF<T, DESCRIPTOR>(e, member, LBMH_FIELD_VELOCITY, "velocity", nullptr) {}
};

/// SuperLatticePhysVelocity3D (latticePhysVelocity3D.hh:84-96)
template <typename T, typename DESCRIPTOR>
class SuperLatticePhysVelocity3D : public detail::LatticeFieldF<T, DESCRIPTOR> {
 public:
  SuperLatticePhysVelocity3D(SuperLattice<T, DESCRIPTOR>& l, const UnitConverter<T, DESCRIPTOR>& c)
      : detail::LatticeFieldF<T, DESCRIPTOR>(l, LBMH_FIELD_PHYS_VELOCITY, "physVelocity",
                                             detail::conversion_of(c).data()) {}
  SuperLatticePhysVelocity3D(Ensemble<T, DESCRIPTOR>& e, int member, const UnitConverter<T, DESCRIPTOR>& c)
      : detail::LatticeFieldF<T, DESCRIPTOR>(e, member, LBMH_FIELD_PHYS_VELOCITY, "physVelocity",
                                             detail::conversion_of(c).data()) {}
};

/// SuperLatticePhysPressure3D (latticePhysPressure3D.hh:71-78)
template <typename T, typename DESCRIPTOR>
class SuperLatticePhysPressure3D : public detail::LatticeFieldF<T, DESCRIPTOR> {
 public:
  SuperLatticePhysPressure3D(SuperLattice<T, DESCRIPTOR>& l, const UnitConverter<T, DESCRIPTOR>& c)
      : detail::LatticeFieldF<T, DESCRIPTOR>(l, LBMH_FIELD_PHYS_PRESSURE, "physPressure",
                                             detail::conversion_of(c).data()) {}
  SuperLatticePhysPressure3D(Ensemble<T, DESCRIPTOR>& e, int member, const UnitConverter<T, DESCRIPTOR>& c)
      : detail::LatticeFieldF<T, DESCRIPTOR>(e, member, LBMH_FIELD_PHYS_PRESSURE, "physPressure",
                                             detail::conversion_of(c).data()) {}
};

/// SuperVTMwriter3D<T> (src/io/superVtmWriter3D.h:50-144) for a one-block
/// lattice: <vtkOutDir><name>.pvd, data/<name>_iT<7>.vtm and
/// data/<name>_iT<7>iC00000.vti with one Float32 DataArray per functor,
/// base64 after a uint32 size prefix, zlib-compressed by default.  The
/// piece's Origin and Spacing are the cuboid's (setOrigin / setDeltaR;
/// 0 and 1 by default, a cuboid decomposition of spacing 1).
template <typename T, typename OUT_T = float>
class SuperVTMwriter3D {
  static_assert(std::is_same_v<OUT_T, float>, "OpenLB writes its VTK buffers as float");

 public:
  SuperVTMwriter3D(const std::string& name, int overlap = 1, bool binary = true, bool compress = true)
      : _name(name), _overlap(overlap), _binary(binary), _compress(compress)
  {
  }

  /// Any of the lattice functors above (of either descriptor).
  template <typename F>
  void addFunctor(F& f)
  {
    _functors.push_back(Entry{[&f]() -> std::string& { return f.getName(); }, f.getTargetDim(),
                              [&f]() { return f.extent(); }, [&f](int overlap) { return f.evaluate(overlap); }});
  }
  template <typename F>
  void addFunctor(F& f, const std::string& name)
  {
    f.getName() = name;
    addFunctor(f);
  }
  void clearAddedFunctors() { _functors.clear(); }
  std::string getName() const { return _name; }
  void setOrigin(std::array<T, 3> origin) { _origin = origin; }
  void setDeltaR(T delta) { _delta = delta; }
  /// singleton::directories().setVtkOutDir: the directory and its data/ must exist
  static void setVtkOutDir(std::string dir) { vtkOutDir() = std::move(dir); }

  void createMasterFile()
  {
    std::ofstream f(vtkOutDir() + _name + ".pvd", std::ios::trunc);
    f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
      << "<Collection>\n</Collection>\n</VTKFile>\n";
    if (!f) throw std::runtime_error("cannot write " + vtkOutDir() + _name + ".pvd");
  }

  void write(int iT)
  {
    if (_functors.empty()) throw std::runtime_error("No functor to write");
    const std::string step = fileName(iT, -1);
    // the master file gets one DataSet line before its closing tags (25 characters)
    const std::string pvd = vtkOutDir() + _name + ".pvd";
    std::string text;
    {
      std::ifstream in(pvd);
      if (!in) throw std::runtime_error("cannot read " + pvd + " (call createMasterFile first)");
      text.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    }
    text.resize(text.size() >= 25 ? text.size() - 25 : 0);
    text += "<DataSet timestep=\"" + std::to_string(iT) + "\" group=\"\" part=\"\" file=\"data/" + step +
            ".vtm\"/>\n</Collection>\n</VTKFile>\n";
    std::ofstream(pvd, std::ios::trunc) << text;
    std::ofstream vtm(vtkOutDir() + "data/" + step + ".vtm", std::ios::trunc);
    vtm << "<?xml version=\"1.0\"?>\n<VTKFile type=\"vtkMultiBlockDataSet\" version=\"1.0\" "
        << "byte_order=\"LittleEndian\">\n<vtkMultiBlockDataSet>\n"
        << "<Block index=\"0\" >\n<DataSet index= \"0\" file=\"" << fileName(iT, 0) << ".vti\">\n</DataSet>\n</Block>\n"
        << "</vtkMultiBlockDataSet>\n</VTKFile>\n";
    if (!vtm) throw std::runtime_error("cannot write " + vtkOutDir() + "data/" + step + ".vtm");
    writeVTI(vtkOutDir() + "data/" + fileName(iT, 0) + ".vti");
  }

 private:
  struct Entry {
    std::function<std::string&()> name;
    int dim;
    std::function<std::array<int, 3>()> extent;
    std::function<std::vector<float>(int)> evaluate;
  };

  static std::string& vtkOutDir()
  {
    static std::string dir = "./tmp/vtkData/";
    return dir;
  }

  std::string fileName(int iT, int iC) const
  {
    char buf[64];
    if (iC < 0) std::snprintf(buf, sizeof(buf), "_iT%07d", iT);
    else std::snprintf(buf, sizeof(buf), "_iT%07diC%05d", iT, iC);
    return _name + buf;
  }

  void writeVTI(const std::string& path)
  {
    const std::array<int, 3> n = _functors.front().extent();
    std::ostringstream out;
    out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"0.1\" ";
    out << (_compress ? "byte_order=\"LittleEndian\" compressor=\"vtkZLibDataCompressor\">\n"
                      : "byte_order=\"LittleEndian\">\n");
    std::ostringstream ext;
    ext << -_overlap << " " << n[0] + _overlap - 1 << " " << -_overlap << " " << n[1] + _overlap - 1 << " "
        << -_overlap << " " << n[2] + _overlap - 1;
    const double d = double(_delta);
    out << "<ImageData WholeExtent=\"" << ext.str() << "\" Origin=\"" << double(_origin[0]) << " "
        << double(_origin[1]) << " " << double(_origin[2]) << "\" Spacing=\"" << d << " " << d << " " << d << "\">\n";
    out << "<Piece Extent=\"" << ext.str() << "\">\n<PointData>\n";
    for (auto& f : _functors) {
      std::string name = f.name();
      for (char& c : name)
        if (c == '<' || c == '>') c = '_';
      out << "<DataArray type=\"Float32\" Name=\"" << name << "\" NumberOfComponents=\"" << f.dim << "\" ";
      out << ((_compress || _binary) ? "format=\"binary\" encoding=\"base64\">\n" : "format=\"ascii\" >\n");
      const std::vector<float> data = f.evaluate(_overlap);
      const auto* bytes = reinterpret_cast<const std::uint8_t*>(data.data());
      const uint32_t size = uint32_t(data.size() * sizeof(float));
      if (_compress) {
        uLongf packed = compressBound(size);
        std::vector<std::uint8_t> z(packed);
        if (compress2(z.data(), &packed, bytes, size, -1) != Z_OK) throw std::runtime_error("zlib compress2 failed");
        const uint32_t head[4] = {1, size, size, uint32_t(packed)};
        out << detail::base64_encode(reinterpret_cast<const std::uint8_t*>(head), sizeof(head))
            << detail::base64_encode(z.data(), packed);
      } else if (_binary) {
        out << detail::base64_encode(reinterpret_cast<const std::uint8_t*>(&size), sizeof(size))
            << detail::base64_encode(bytes, size);
      } else {
        for (float v : data) out << v << " ";
      }
      out << "\n</DataArray>\n";
    }
    out << "</PointData>\n</Piece>\n</ImageData>\n</VTKFile>\n";
    std::ofstream file(path, std::ios::trunc);
    file << out.str();
    if (!file) throw std::runtime_error("cannot write " + path);
  }

  std::string _name;
  int _overlap;
  bool _binary, _compress;
  std::array<T, 3> _origin{T(0), T(0), T(0)};
  T _delta = T(1);
  std::vector<Entry> _functors;
};

}  // namespace hip
}  // namespace olb
