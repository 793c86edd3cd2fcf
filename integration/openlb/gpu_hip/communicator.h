/* Overlap communication of Platform::GPU_HIP blocks on one device.
 *
 * The generic block communicator copies the requested fields of the cells a
 * BlockCommunicationNeighborhood lists, through host buffers
 * (src/core/blockLattice.hh:752-870).  Between GPU_HIP blocks the populations
 * are copied device to device instead: one lbmh_halo per neighbour, built
 * from the same cell lists once both blocks have device lattices, and run by
 * lbmh_halo_copy on every communicate() (the stage order of
 * SuperLattice::collideAndStream puts it after all blocks collided and before
 * they stream, src/core/superLattice.hh:629-637).  Only populations cross this
 * way; a neighbourhood that requests other fields of a GPU_HIP block is
 * refused.  Builds without MPI only.
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
  };

  std::vector<Task> _tasks;

  /// padded cell id of a block with overlap p as the lattice's (overlap 1) id
  static std::uint32_t latticeCell(const Block& block, CellID iCell)
  {
    const auto r = block.getLatticeR(iCell);
    for (unsigned d = 0; d < 3; ++d) {
      if (r[d] < -1 || r[d] > block.getExtent()[d]) {
        throw std::runtime_error("GPU_HIP: overlap communication beyond the first overlap layer");
      }
    }
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
    std::vector<std::uint32_t> dst, src;
    dst.reserve(task.targetCells->size());
    src.reserve(task.sourceCells->size());
    for (CellID iCell : *task.targetCells) {
      dst.push_back(latticeCell(*task.target, iCell));
    }
    for (CellID iCell : *task.sourceCells) {
      src.push_back(latticeCell(*task.source, iCell));
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
      for (const std::type_index& field : neighborhood.getFieldsCommonWith(localC)) {
        if (field != typeid(descriptors::POPULATION)) {
          throw std::logic_error("GPU_HIP: blocks exchange populations only");
        }
      }
      if (loadBalancer.platform(loadBalancer.loc(localC)) != Platform::GPU_HIP) {
        throw std::logic_error("GPU_HIP: overlap communication with a block of another platform");
      }
      Task task;
      task.target = &super.template getBlock<Block>(iC);
      task.source = &super.template getBlock<Block>(loadBalancer.loc(localC));
      task.targetCells = &neighborhood.getCellsInboundFrom(localC);
      task.sourceCells = &neighborhood.getCellsRequestedFrom(localC);
      _tasks.push_back(task);
    });
  }

  ~ConcreteBlockCommunicator()
  {
    for (auto& task : _tasks) {
      if (task.halo) {
        lbmh_halo_destroy(task.halo);
      }
    }
  }

  void copy() override
  {
    for (auto& task : _tasks) {
      build(task);
      gpu::hip::check(lbmh_halo_copy(task.halo, nullptr), "lbmh_halo_copy");
    }
  }

};

}

#endif
