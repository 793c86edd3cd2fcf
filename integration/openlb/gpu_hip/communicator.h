/* Overlap communication of Platform::GPU_HIP blocks on one device.
 *
 * The generic block communicator copies the requested fields of the cells a
 * BlockCommunicationNeighborhood lists, through host buffers
 * (src/core/blockLattice.hh:752-870).  Between GPU_HIP blocks the populations
 * are copied device to device instead: one lbmh_halo per neighbour, built
 * from the same cell lists once both blocks have device lattices, and run by
 * lbmh_halo_copy on every communicate() (the stage order of
 * SuperLattice::collideAndStream puts it after all blocks collided and before
 * they stream, src/core/superLattice.hh:629-637).  The device lattice holds
 * one overlap layer; cells of a deeper exchange (the full one of
 * SuperLattice::communicate, overlap 2) are copied between the host columns
 * as well.  The halos also tell each block which block owns the neighbours
 * its boundary post-processors read (lbmh_set_overlap_ids), so every halo is
 * built before the blocks' first post-processing.  Only populations cross
 * this way: other fields a neighbourhood requests (FORCE of the full
 * exchange) keep the overlap values the set-up wrote, which is all the device
 * needs, since it reads them at core cells.  Builds without MPI only.
 */
#ifndef OLB_GPU_HIP_COMMUNICATOR_H
#define OLB_GPU_HIP_COMMUNICATOR_H

#ifdef PARALLEL_MODE_MPI
#error "the GPU_HIP platform covers one process; build without PARALLEL_MODE_MPI"
#endif

#include <typeindex>

namespace olb {

template <typename T, typename DESCRIPTOR>
class ConcreteBlockCommunicator<ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>> final
  : public BlockCommunicator {
private:
  using Block = ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>;

  struct Task {
    Block* target;
    Block* source;
    const std::vector<CellID>* targetCells;
    const std::vector<CellID>* sourceCells;
    lbmh_halo* halo = nullptr;
    lbmh_lattice* targetLattice = nullptr;
    lbmh_lattice* sourceLattice = nullptr;
    bool deep = false;  // some target cells lie beyond the first overlap layer
  };

  std::vector<Task> _tasks;

  static bool firstLayer(const Block& block, CellID iCell)
  {
    const auto r = block.getLatticeR(iCell);
    for (unsigned d = 0; d < 3; ++d) {
      if (r[d] < -1 || r[d] > block.getExtent()[d]) {
        return false;
      }
    }
    return true;
  }

  /// padded cell id of a block with overlap p as the lattice's (overlap 1) id
  static std::uint32_t latticeCell(const Block& block, CellID iCell)
  {
    const auto r = block.getLatticeR(iCell);
    const std::uint64_t sy = std::uint64_t(block.getNz()) + 2;
    const std::uint64_t sx = (std::uint64_t(block.getNy()) + 2) * sy;
    return std::uint32_t((r[0] + 1) * sx + (r[1] + 1) * sy + (r[2] + 1));
  }

  void build(Task& task)
  {
    auto target = gpu::hip::stateOf(*task.target);
    auto source = gpu::hip::stateOf(*task.source);
    lbmh_lattice* tl = target->lattice();
    lbmh_lattice* sl = source->lattice();
    if (task.halo && tl == task.targetLattice && sl == task.sourceLattice) {
      return;
    }
    if (task.halo) {
      lbmh_halo_destroy(task.halo);
      task.halo = nullptr;
    }
    // the device lattice holds the first overlap layer: its cells get a halo
    std::vector<std::uint32_t> dst, src;
    dst.reserve(task.targetCells->size());
    src.reserve(task.sourceCells->size());
    for (std::size_t k = 0; k < task.targetCells->size(); ++k) {
      const CellID to = (*task.targetCells)[k], from = (*task.sourceCells)[k];
      if (!firstLayer(*task.target, to)) {
        continue;
      }
      if (!firstLayer(*task.source, from)) {
        throw std::runtime_error("GPU_HIP: overlap cell copied from beyond the source block's first layer");
      }
      dst.push_back(latticeCell(*task.target, to));
      src.push_back(latticeCell(*task.source, from));
    }
    gpu::hip::check(lbmh_halo_create(tl, sl, dst.data(), src.data(), dst.size(), &task.halo), "lbmh_halo_create");
    task.targetLattice = tl;
    task.sourceLattice = sl;
  }

public:
  ConcreteBlockCommunicator(SuperLattice<T,DESCRIPTOR>& super,
                            LoadBalancer<T>& loadBalancer,
                            int iC,
                            const BlockCommunicationNeighborhood<T,DESCRIPTOR::d>& neighborhood)
  {
    neighborhood.forNeighbors([&](int localC) {
      if (neighborhood.getCellsInboundFrom(localC).empty()) {
        return;
      }
      // only the populations cross: the device reads the other fields (FORCE)
      // at core cells only, and OpenLB's set-up (defineField over a material
      // indicator) writes their overlap copies too
      bool populations = false;
      for (const std::type_index& field : neighborhood.getFieldsCommonWith(localC)) {
        populations = populations || field == typeid(descriptors::POPULATION);
      }
      if (!populations) {
        return;
      }
      if (loadBalancer.platform(loadBalancer.loc(localC)) != Platform::GPU_HIP) {
        throw std::logic_error("GPU_HIP: overlap communication with a block of another platform");
      }
      Task task;
      task.target = &super.template getBlock<Block>(iC);
      task.source = &super.template getBlock<Block>(loadBalancer.loc(localC));
      task.targetCells = &neighborhood.getCellsInboundFrom(localC);
      task.sourceCells = &neighborhood.getCellsRequestedFrom(localC);
      if (task.targetCells->size() != task.sourceCells->size()) {
        throw std::logic_error("GPU_HIP: overlap cell lists of different lengths");
      }
      for (CellID iCell : *task.targetCells) {
        task.deep = task.deep || !firstLayer(*task.target, iCell);
      }
      _tasks.push_back(task);
    });
    if (!_tasks.empty()) {
      gpu::hip::HaloBuilders<T,DESCRIPTOR>::all()[this] = [this]() {
        for (auto& task : _tasks) {
          build(task);
        }
      };
    }
  }

  ~ConcreteBlockCommunicator()
  {
    gpu::hip::HaloBuilders<T,DESCRIPTOR>::all().erase(this);
    for (auto& task : _tasks) {
      if (task.halo) {
        lbmh_halo_destroy(task.halo);
      }
    }
  }

  /// The full overlap exchange (SuperLattice::communicate, overlap 2, before
  /// functors and output read the host columns) reaches beyond the device
  /// lattice's layer: such an exchange is also done between the host columns,
  /// which the switch to ProcessingContext::Evaluation filled from the device.
  void copy() override
  {
    // this block's neighbours in one launch
    std::vector<lbmh_halo*> halos;
    halos.reserve(_tasks.size());
    for (auto& task : _tasks) {
      build(task);
      halos.push_back(task.halo);
    }
    gpu::hip::check(lbmh_halo_copy_many(halos.data(), halos.size(), nullptr), "lbmh_halo_copy_many");
    for (auto& task : _tasks) {
      if (task.deep) {
        auto& to = task.target->template getField<descriptors::POPULATION>();
        auto& from = task.source->template getField<descriptors::POPULATION>();
        for (unsigned iPop = 0; iPop < DESCRIPTOR::q; ++iPop) {
          for (std::size_t k = 0; k < task.targetCells->size(); ++k) {
            to[iPop][(*task.targetCells)[k]] = from[iPop][(*task.sourceCells)[k]];
          }
        }
      }
    }
  }

};

}

#endif
