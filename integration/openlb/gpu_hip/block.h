/* Device state of one Platform::GPU_HIP block: the liblbmh lattice that holds
 * its populations in HBM, and the translation of the block's dynamics,
 * parameters and boundary fields into the lattice's set-up.
 *
 * OpenLB decides per cell which dynamics runs (BlockDynamicsMap) and keeps
 * the values those dynamics read in per-cell fields.  The lattice runs one
 * bulk collision, a dynamics-id byte per cell, a table of moving walls and
 * the local velocity boundaries (include/lbmh.h).  BlockState builds that
 * set-up from the block whenever the assignment of dynamics changed or the
 * block switched to ProcessingContext::Simulation (after which the host may
 * have changed field values), and drives the time step:
 *   collide   ConcreteBlockLattice::collide         -> lbmh_collide
 *   stream    ConcreteBlockLattice::stream          -> lbmh_stream (shift +
 *             the boundary post-processors of the PostStream stage)
 *   PostStream post-processors outside a time step  -> lbmh_initialize
 * (src/core/superLattice.hh:601-663, src/core/blockLattice.hh:56-67, 303-339).
 */
#ifndef OLB_GPU_HIP_BLOCK_H
#define OLB_GPU_HIP_BLOCK_H

#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace olb {

namespace gpu::hip {

/// The device halos of the process's GPU_HIP block communicators, by
/// communicator.  A block's first post-processing (BlockLattice::initialize)
/// comes before any communication; its post-processors that read cells of
/// other blocks find them through the halos, so every halo is built first.
template <typename T, typename DESCRIPTOR>
struct HaloBuilders {
  static std::map<const void*, std::function<void()>>& all()
  {
    static std::map<const void*, std::function<void()>> builders;
    return builders;
  }

  static void buildAll()
  {
    for (auto& entry : all()) {
      entry.second();
    }
  }
};

template <typename T, typename DESCRIPTOR>
class BlockState final : public PopulationMirror {
public:
  using Block = ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>;

  /// A collision operator of the block: its dynamics type's role and cells
  struct DynamicsEntry {
    RoleInfo info;
    const Dynamics<T,DESCRIPTOR>* dynamics;
    std::function<void(lbmh_bulk_params&)> parameters;  // bulk roles: OMEGA, MAGIC, SMAGORINSKY
    std::string name;
  };

  /// A cell-wise post-processor of the block and its cells
  struct PostEntry {
    PostInfo info;
    const std::vector<CellID>* cells;
    std::string name;
  };

private:
  Block& _block;
  lbmh_lattice* _lattice = nullptr;
  int _kind = -1;

  std::vector<DynamicsEntry> _dynamics;
  std::vector<PostEntry> _posts;

  bool _structureChanged = true;   // dynamics assignment or operators changed
  bool _valuesChanged = true;      // host field values may have changed
  std::vector<bool> _uploaded;     // columns uploaded since the last set-up

  lbmh_bulk_params _params{};
  std::vector<double> _forceField;  // per-cell FORCE [3][core cells] on the device, empty for a uniform one
  std::array<double,4> _statistics{0.0, 1.0, 0.0, 0.0};
  bool _collided = false;
  bool _postDone = false;

  static constexpr int q = DESCRIPTOR::q;

  int coreIndex(int x, int y, int z) const
  {
    return (x * _block.getNy() + y) * _block.getNz() + z;
  }

  static std::string vec(const std::array<int,3>& n)
  {
    std::ostringstream s;
    s << "(" << n[0] << ", " << n[1] << ", " << n[2] << ")";
    return s.str();
  }

  /// (Re)create the lattice for bulk kind `kind`, keeping the populations
  void ensureLattice(int kind)
  {
    if (_lattice && kind == _kind) {
      return;
    }
    if (_lattice) {
      // carry the populations over through the host columns
      auto& population = _block.template getField<descriptors::POPULATION>();
      for (unsigned iPop = 0; iPop < q; ++iPop) {
        check(lbmh_column_download(_lattice, int(iPop), population[iPop].data(), population[iPop].getShift(),
                                   _block.getPadding(), hostType<T>()), "lbmh_column_download");
      }
      lbmh_lattice_destroy(_lattice);
      _lattice = nullptr;
    }
    lbmh_lattice_desc desc{};
    desc.q = q;
    desc.nx = _block.getNx();
    desc.ny = _block.getNy();
    desc.nz = _block.getNz();
    // overlap cells are filled by the block communicators, periodic ones included
    desc.periodic[0] = desc.periodic[1] = desc.periodic[2] = 0;
    desc.storage = hostType<T>();
    desc.bulk_kind = kind;
    desc.device = 0;
    check(lbmh_lattice_create(&desc, &_lattice), "lbmh_lattice_create");
    _kind = kind;
    // the new lattice holds no parameters yet: all-ones bytes (NaNs) match no block's
    std::memset(&_params, 0xff, sizeof(_params));
    _forceField.clear();
    std::fill(_uploaded.begin(), _uploaded.end(), false);
    auto& population = _block.template getField<descriptors::POPULATION>();
    for (unsigned iPop = 0; iPop < q; ++iPop) {
      check(lbmh_column_upload(_lattice, int(iPop), population[iPop].data(), population[iPop].getShift(),
                               _block.getPadding(), hostType<T>()), "lbmh_column_upload");
    }
    _structureChanged = _valuesChanged = true;
  }

  /// The lattice's dynamics ids, walls and local boundaries from the block
  void configure()
  {
    const int nx = _block.getNx(), ny = _block.getNy(), nz = _block.getNz();
    const std::size_t ncore = std::size_t(nx) * ny * nz;

    std::map<const Dynamics<T,DESCRIPTOR>*, const DynamicsEntry*> byDynamics;
    for (const auto& d : _dynamics) {
      byDynamics[d.dynamics] = &d;
    }
    std::vector<const DynamicsEntry*> role(ncore, nullptr);
    const DynamicsEntry* bulk = nullptr;
    for (int x = 0; x < nx; ++x) {
      for (int y = 0; y < ny; ++y) {
        for (int z = 0; z < nz; ++z) {
          const Dynamics<T,DESCRIPTOR>* dyn = _block.getDynamics(_block.getCellId(x, y, z));
          auto it = byDynamics.find(dyn);
          if (it == byDynamics.end()) {
            throw std::runtime_error("GPU_HIP: cell (" + std::to_string(x) + ", " + std::to_string(y) + ", " +
                                     std::to_string(z) + ") has dynamics without a GPU_HIP collision operator");
          }
          const DynamicsEntry* e = it->second;
          if (e->info.role == Role::Unsupported) {
            throw std::runtime_error("GPU_HIP: dynamics " + e->name + " is not available on this platform");
          }
          if (e->info.role == Role::Bulk) {
            if (bulk && bulk != e) {
              throw std::runtime_error("GPU_HIP: one block runs one bulk dynamics, this one has " + bulk->name +
                                       " and " + e->name);
            }
            bulk = e;
          }
          role[std::size_t(coreIndex(x, y, z))] = e;
        }
      }
    }

    // walls and local boundaries, grouped by their values
    std::vector<std::uint8_t> ids(ncore, 0);
    std::map<std::array<T,4>, int> walls;     // (rho, u) -> id
    std::map<std::array<T,3>, int> velocity;  // u -> material id of a velocity group (faces, edges, corners)
    std::map<int, int> velocityKind;          // velocity group -> LBMH_LOCAL_VELOCITY / LBMH_INTERP_VELOCITY
    // (rho, u with its normal component 0) -> material id; [0] LocalPressure, [1] InterpolatedPressure faces
    std::map<std::array<T,4>, int> pressure[2];
    int next = LBMH_ID_FIRST_WALL;
    auto take = [&]() {
      if (next > 200) {
        throw std::runtime_error("GPU_HIP: more distinct wall / boundary values than dynamics ids");
      }
      return next++;
    };
    bool local = false;
    for (int x = 0; x < nx; ++x) {
      for (int y = 0; y < ny; ++y) {
        for (int z = 0; z < nz; ++z) {
          const std::size_t c = std::size_t(coreIndex(x, y, z));
          const Role r = role[c]->info.role;
          switch (r) {
          case Role::None:
            ids[c] = LBMH_ID_NONE;
            break;
          case Role::Bulk:
            ids[c] = LBMH_ID_BULK;
            break;
          case Role::BounceBack:
            ids[c] = LBMH_ID_BOUNCE_BACK;
            break;
          case Role::MovingWall: {
            auto cell = _block.get(x, y, z);
            T rho = cell.computeRho();
            T u[3] = {};
            cell.computeU(u);
            const std::array<T,4> key{rho, u[0], u[1], u[2]};
            auto it = walls.find(key);
            if (it == walls.end()) {
              it = walls.emplace(key, take()).first;
            }
            ids[c] = std::uint8_t(it->second);
            break;
          }
          case Role::PressureFlat:
          case Role::InterpPressureFlat: {
            // fixed density and the tangential velocity; the normal component
            // comes from the populations every step
            auto cell = _block.get(x, y, z);
            T u[3] = {};
            cell.computeU(u);
            const auto& nrm = role[c]->info.normal;
            for (int d = 0; d < 3; ++d) {
              if (nrm[d] != 0) {
                u[d] = T(0);
              }
            }
            const std::array<T,4> key{cell.computeRho(), u[0], u[1], u[2]};
            auto& groups = pressure[r == Role::InterpPressureFlat ? 1 : 0];
            auto it = groups.find(key);
            if (it == groups.end()) {
              it = groups.emplace(key, take()).first;
            }
            ids[c] = std::uint8_t(it->second);
            local = true;
            break;
          }
          default: {  // the velocity family: faces, outer / inner edges and corners
            auto cell = _block.get(x, y, z);
            T u[3] = {};
            cell.computeU(u);
            const std::array<T,3> key{u[0], u[1], u[2]};
            auto it = velocity.find(key);
            if (it == velocity.end()) {
              it = velocity.emplace(key, take()).first;
            }
            ids[c] = std::uint8_t(it->second);
            if (r == Role::VelocityFlat || r == Role::InterpVelocityFlat) {
              // the faces decide whether the group is local or interpolated
              const int kind = r == Role::InterpVelocityFlat ? LBMH_INTERP_VELOCITY : LBMH_LOCAL_VELOCITY;
              auto k = velocityKind.emplace(it->second, kind).first;
              if (k->second != kind) {
                throw std::runtime_error("GPU_HIP: LocalVelocity and InterpolatedVelocity faces with the same "
                                         "velocity share edges; give them different velocities or blocks");
              }
            }
            local = true;
            break;
          }
          }
        }
      }
    }

    // the first overlap layer as the classification sees it: the materials of
    // the neighbouring blocks' cells (or periodic images), 0 beyond the domain
    const std::vector<std::uint8_t> ring = overlapMaterials(byDynamics);
    if (local) {
      crossCheckBoundaries(ids, ring, role);
    }
    // everything above is host work that validates the block; the device comes last
    ensureLattice(bulk ? bulk->info.bulkKind : LBMH_BGK);
    check(lbmh_set_dynamics_ids(_lattice, ids.data()), "lbmh_set_dynamics_ids");
    check(lbmh_set_overlap_ids(_lattice, ring.data()), "lbmh_set_overlap_ids");
    for (const auto& [key, id] : walls) {
      const double u[3] = {double(key[1]), double(key[2]), double(key[3])};
      check(lbmh_set_wall(_lattice, id, double(key[0]), u), "lbmh_set_wall");
    }
    for (const auto& [key, id] : velocity) {
      const double u[3] = {double(key[0]), double(key[1]), double(key[2])};
      auto k = velocityKind.find(id);
      const int kind = k == velocityKind.end() ? LBMH_LOCAL_VELOCITY : k->second;
      check(lbmh_set_local_boundary(_lattice, id, kind, 1.0, u), "lbmh_set_local_boundary");
    }
    for (int interp = 0; interp < 2; ++interp) {
      for (const auto& [key, id] : pressure[interp]) {
        const double u[3] = {double(key[1]), double(key[2]), double(key[3])};
        check(lbmh_set_local_boundary(_lattice, id, interp ? LBMH_INTERP_PRESSURE : LBMH_LOCAL_PRESSURE,
                                      double(key[0]), u), "lbmh_set_local_boundary");
      }
    }
    readParameters(bulk, role);
    _structureChanged = _valuesChanged = false;
    std::fill(_uploaded.begin(), _uploaded.end(), false);
  }

  /// Materials of the first overlap layer, padded [nx+2][ny+2][nz+2] (core
  /// entries 0): OpenLB defines dynamics on the overlap cells as on the core
  /// ones (defineDynamics and boundary::set visit every spatial location,
  /// src/core/blockLattice.hh:277-301, src/boundary/setBoundary.h:83-101), so
  /// a cell's collision operator tells fluid (1), outside (0) and wall (2)
  /// apart there as in the core
  std::vector<std::uint8_t> overlapMaterials(
    const std::map<const Dynamics<T,DESCRIPTOR>*, const DynamicsEntry*>& byDynamics) const
  {
    const int nx = _block.getNx(), ny = _block.getNy(), nz = _block.getNz();
    std::vector<std::uint8_t> ring(std::size_t(nx + 2) * (ny + 2) * (nz + 2), 0);
    for (int x = -1; x <= nx; ++x) {
      for (int y = -1; y <= ny; ++y) {
        for (int z = -1; z <= nz; ++z) {
          if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz) {
            continue;
          }
          auto it = byDynamics.find(_block.getDynamics(_block.getCellId(x, y, z)));
          const Role r = it == byDynamics.end() ? Role::Unsupported : it->second->info.role;
          const std::uint8_t m = r == Role::None ? LBMH_ID_NONE : (r == Role::Bulk ? LBMH_ID_BULK : LBMH_ID_BOUNCE_BACK);
          ring[(std::size_t(x + 1) * (ny + 2) + std::size_t(y + 1)) * (nz + 2) + std::size_t(z + 1)] = m;
        }
      }
    }
    return ring;
  }

  /// The library classifies local-boundary cells from the id map (and the
  /// overlap materials) as OpenLB's computeBoundaryTypeAndNormal does; every
  /// cell must come out with the type and normal its OpenLB dynamics (and
  /// post-processor) were chosen for.
  void crossCheckBoundaries(const std::vector<std::uint8_t>& ids, const std::vector<std::uint8_t>& ring,
                            const std::vector<const DynamicsEntry*>& role)
  {
    const int nx = _block.getNx(), ny = _block.getNy(), nz = _block.getNz();
    const std::size_t ncore = ids.size();
    const int extent[3] = {nx, ny, nz};
    const int periodic[3] = {0, 0, 0};
    std::vector<std::int8_t> cls(4 * ncore);
    check(lbmh_classify_boundary_overlap(extent, periodic, ids.data(), ring.data(), cls.data()),
          "lbmh_classify_boundary_overlap");
    std::map<CellID, const PostEntry*> post;
    for (const auto& p : _posts) {
      for (CellID iCell : *p.cells) {
        post[iCell] = &p;
      }
    }
    for (int x = 0; x < nx; ++x) {
      for (int y = 0; y < ny; ++y) {
        for (int z = 0; z < nz; ++z) {
          const std::size_t c = std::size_t(coreIndex(x, y, z));
          const RoleInfo& info = role[c]->info;
          if (info.role == Role::None || info.role == Role::Bulk || info.role == Role::BounceBack ||
              info.role == Role::MovingWall) {
            continue;
          }
          const int type = cls[c];
          const std::array<int,3> n{cls[ncore + c], cls[2 * ncore + c], cls[3 * ncore + c]};
          int want = -1;
          std::array<int,3> normal = info.normal;
          if (info.role == Role::VelocityFlat || info.role == Role::PressureFlat ||
              info.role == Role::InterpVelocityFlat || info.role == Role::InterpPressureFlat) {
            want = kFlat;
          } else if (info.role == Role::VelocityInnerCorner) {
            want = kInternalCorner;
          } else if (info.role == Role::VelocityInnerEdge) {
            want = kInternalEdge;
          } else {
            auto it = post.find(_block.getCellId(x, y, z));
            if (it == post.end()) {
              throw std::runtime_error("GPU_HIP: fixed-velocity boundary cell without its edge / corner "
                                       "post-processor at (" + std::to_string(x) + ", " + std::to_string(y) +
                                       ", " + std::to_string(z) + "); the block has " +
                                       std::to_string(_posts.size()) + " such post-processors on " +
                                       std::to_string(post.size()) + " cells");
            }
            want = it->second->info.type;
            normal = it->second->info.normal;
          }
          if (type != want || n != normal) {
            throw std::runtime_error("GPU_HIP: boundary cell (" + std::to_string(x) + ", " + std::to_string(y) +
                                     ", " + std::to_string(z) + ") is set up as type " + std::to_string(want) +
                                     " normal " + vec(normal) + ", the device classifies it as type " +
                                     std::to_string(type) + " normal " + vec(n));
          }
        }
      }
    }
  }

  /// Bulk parameters; the forced kinds read the FORCE field: one value for all
  /// bulk cells goes into the parameters, a varying one to the device field
  void readParameters(const DynamicsEntry* bulk, const std::vector<const DynamicsEntry*>& role)
  {
    lbmh_bulk_params p{};
    p.omega = 1.0;
    if (bulk && bulk->parameters) {
      bulk->parameters(p);
    }
    std::vector<double> field;
    if (_kind == LBMH_FORCED_BGK || _kind == LBMH_FORCED_TRT) {
      if constexpr (DESCRIPTOR::template provides<descriptors::FORCE>()) {
        auto& force = _block.template getField<descriptors::FORCE>();
        const int nx = _block.getNx(), ny = _block.getNy(), nz = _block.getNz();
        const std::size_t ncore = std::size_t(nx) * ny * nz;
        field.assign(3 * ncore, 0.0);
        bool first = true;
        bool uniform = true;
        for (int x = 0; x < nx; ++x) {
          for (int y = 0; y < ny; ++y) {
            for (int z = 0; z < nz; ++z) {
              const std::size_t c = std::size_t(coreIndex(x, y, z));
              if (role[c]->info.role != Role::Bulk) {
                continue;
              }
              const CellID iCell = _block.getCellId(x, y, z);
              for (int d = 0; d < 3; ++d) {
                const double f = double(force[d][iCell]);
                field[d * ncore + c] = f;
                if (first) {
                  p.force[d] = f;
                } else if (f != p.force[d]) {
                  uniform = false;
                }
              }
              first = false;
            }
          }
        }
        if (uniform) {
          field.clear();
        }
      } else {
        throw std::runtime_error("GPU_HIP: forced dynamics on a descriptor without FORCE");
      }
    }
    if (std::memcmp(&p, &_params, sizeof(p)) != 0) {
      check(lbmh_set_bulk_params(_lattice, &p), "lbmh_set_bulk_params");
      _params = p;
    }
    if (field != _forceField) {
      check(lbmh_set_force_field(_lattice, field.empty() ? nullptr : field.data()), "lbmh_set_force_field");
      _forceField = std::move(field);
    }
  }

  void prepare()
  {
    if (_structureChanged || _valuesChanged || !_lattice) {
      configure();
    } else {
      // parameters may change between steps (setParameter); they are cheap to compare
      const DynamicsEntry* bulk = nullptr;
      for (const auto& d : _dynamics) {
        if (d.info.role == Role::Bulk && d.parameters) {
          bulk = &d;
        }
      }
      if (bulk && _kind != LBMH_FORCED_BGK && _kind != LBMH_FORCED_TRT) {
        lbmh_bulk_params p = _params;
        bulk->parameters(p);
        if (std::memcmp(&p, &_params, sizeof(p)) != 0) {
          check(lbmh_set_bulk_params(_lattice, &p), "lbmh_set_bulk_params");
          _params = p;
        }
      }
    }
  }

public:
  explicit BlockState(Block& block):
    _block(block),
    _uploaded(q, false)
  {
    auto& population = _block.template getField<descriptors::POPULATION>();
    for (unsigned iPop = 0; iPop < q; ++iPop) {
      population[iPop].attach(this, iPop);
    }
  }

  ~BlockState()
  {
    auto& population = _block.template getField<descriptors::POPULATION>();
    for (unsigned iPop = 0; iPop < q; ++iPop) {
      population[iPop].attach(nullptr, iPop);
    }
    if (_lattice) {
      lbmh_lattice_destroy(_lattice);
    }
  }

  BlockState(const BlockState&) = delete;
  BlockState& operator=(const BlockState&) = delete;

  void add(DynamicsEntry entry)
  {
    _dynamics.push_back(std::move(entry));
    _structureChanged = true;
  }

  void add(PostEntry entry)
  {
    _posts.push_back(std::move(entry));
    _structureChanged = true;
  }

  void markStructure()
  {
    _structureChanged = true;
  }

  lbmh_lattice* lattice()
  {
    prepare();
    return _lattice;
  }

  void upload(unsigned iPop, const void* data, std::ptrdiff_t shift) override
  {
    if (_uploaded[iPop]) {      // a new switch to Simulation: re-read the host's field values
      std::fill(_uploaded.begin(), _uploaded.end(), false);
      _valuesChanged = true;
    }
    if (_valuesChanged || _structureChanged || !_lattice) {
      if (_dynamics.empty()) {
        return;                 // nothing to simulate yet: the block has no collision operators
      }
      configure();              // uploads every column if it creates the lattice
    }
    check(lbmh_column_upload(_lattice, int(iPop), data, shift, _block.getPadding(), hostType<T>()),
          "lbmh_column_upload");
    _uploaded[iPop] = true;
  }

  void download(unsigned iPop, void* data, std::ptrdiff_t shift) override
  {
    if (!_lattice) {
      return;
    }
    check(lbmh_column_download(_lattice, int(iPop), data, shift, _block.getPadding(), hostType<T>()),
          "lbmh_column_download");
  }

  /// ConcreteBlockLattice::collide: once per time step, whichever operator
  /// asks first.  With statistics on, the statistics of the populations the
  /// collision starts from are reduced first, as OpenLB gathers them during
  /// the collision.
  void collide(bool statistics)
  {
    if (_collided) {
      return;
    }
    prepare();
    if (statistics) {
      check(lbmh_statistics(_lattice, _statistics.data()), "lbmh_statistics");
    }
    check(lbmh_collide(_lattice, nullptr), "lbmh_collide");
    _collided = true;
    _postDone = false;
  }

  /// {cells, average rho, average energy, max |u|} of the last collision
  const std::array<double,4>& statistics() const
  {
    return _statistics;
  }

  /// ConcreteBlockLattice::stream
  void stream()
  {
    prepare();
    check(lbmh_stream(_lattice, nullptr), "lbmh_stream");
    _collided = false;
    _postDone = true;
  }

  /// A post-processor the device runs itself: inside a time step lbmh_stream
  /// ran it already, outside (BlockLattice::initialize) it runs once now --
  /// for blocks whose post-processors read one another, once the last of
  /// them got here (lbmh_initialize), through halos built now
  void postStream()
  {
    if (_postDone) {
      return;
    }
    prepare();
    HaloBuilders<T,DESCRIPTOR>::buildAll();
    check(lbmh_initialize(_lattice), "lbmh_initialize");
    _postDone = true;
  }

};

/// The device state of a GPU_HIP block, shared by the block's operators
template <typename T, typename DESCRIPTOR>
std::shared_ptr<BlockState<T,DESCRIPTOR>> stateOf(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>& block)
{
  static std::map<const void*, std::weak_ptr<BlockState<T,DESCRIPTOR>>> states;
  auto& slot = states[&block];
  if (auto state = slot.lock()) {
    return state;
  }
  auto state = std::make_shared<BlockState<T,DESCRIPTOR>>(block);
  slot = state;
  return state;
}

/// ConcreteBlockLattice<...,GPU_HIP>::stream
template <typename T, typename DESCRIPTOR>
void stream(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>& block)
{
  stateOf(block)->stream();
}

/// Synchronisation before a post-processing stage (src/core/superLattice.hh:666-671)
inline void synchronize()
{
  int count = 0;
  if (lbmh_device_count(&count) == LBMH_OK && count > 0) {
    check(lbmh_device_synchronize(0), "lbmh_device_synchronize");
  }
}

}

}

#endif
