/* Collision and post-processing operators of Platform::GPU_HIP blocks.
 *
 * OpenLB instantiates one collision operator per dynamics type of a block and
 * one block operator per post-processor type (src/core/operator.h:43-100,
 * BlockDynamicsMap, BlockPostProcessorMap).  On GPU_HIP each operator keeps
 * the host side of its contract (the cells it covers, the dynamics object for
 * host-side cell access) and registers its role with the block's device state
 * (block.h); the fused device step then covers all of them at once:
 * apply() of whichever collision operator comes first launches the block's
 * collision, and post-processors the device runs are no-ops inside a step.
 */
#ifndef OLB_GPU_HIP_OPERATOR_H
#define OLB_GPU_HIP_OPERATOR_H

#include <typeinfo>

namespace olb {

namespace gpu::hip {

/// Bulk parameters of DYNAMICS read from the block's operator parameters
template <typename T, typename DESCRIPTOR, typename DYNAMICS>
std::function<void(lbmh_bulk_params&)> parameterReader(ParametersOfOperatorD<T,DESCRIPTOR,DYNAMICS>* parameters)
{
  return [parameters](lbmh_bulk_params& p) {
    using list = typename DYNAMICS::parameters;
    if constexpr (list::template contains<descriptors::OMEGA>()) {
      p.omega = double(parameters->template get<descriptors::OMEGA>());
    }
    if constexpr (list::template contains<collision::TRT::MAGIC>()) {
      p.magic = double(parameters->template get<collision::TRT::MAGIC>());
    }
    if constexpr (list::template contains<collision::LES::SMAGORINSKY>()) {
      p.smago = double(parameters->template get<collision::LES::SMAGORINSKY>());
    }
  };
}

}

template <typename T, typename DESCRIPTOR, typename DYNAMICS>
class ConcreteBlockCollisionO<T,DESCRIPTOR,Platform::GPU_HIP,DYNAMICS> final
  : public BlockCollisionO<T,DESCRIPTOR,Platform::GPU_HIP> {
private:
  std::unique_ptr<DYNAMICS> _dynamics;
  ParametersOfOperatorD<T,DESCRIPTOR,DYNAMICS>* _parameters;
  ConcreteBlockMask<T,Platform::GPU_HIP>* _mask;
  std::shared_ptr<gpu::hip::BlockState<T,DESCRIPTOR>> _state;

public:
  ConcreteBlockCollisionO():
    _dynamics(new DYNAMICS()),
    _parameters(nullptr),
    _mask(nullptr)
  { }

  std::type_index id() const override
  {
    return typeid(DYNAMICS);
  }

  std::size_t weight() const override
  {
    return _mask->weight();
  }

  void set(CellID iCell, bool state, bool overlap) override
  {
    if (!overlap) {
      _mask->set(iCell, state);
    }
    if (_state) {
      _state->markStructure();
    }
  }

  Dynamics<T,DESCRIPTOR>* getDynamics() override
  {
    return _dynamics.get();
  }

  void setup(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>& block) override
  {
    _parameters = &block.template getData<OperatorParameters<DYNAMICS>>().parameters;
    _mask = &block.template getData<DynamicsMask<DYNAMICS>>();
    if constexpr (dynamics::has_parametrized_momenta_v<DYNAMICS>) {
      _dynamics->setMomentaParameters(_parameters);
    }
    _state = gpu::hip::stateOf(block);
    constexpr gpu::hip::RoleInfo info = gpu::hip::roleOf<T,DESCRIPTOR,DYNAMICS>();
    std::function<void(lbmh_bulk_params&)> reader;
    if constexpr (info.role == gpu::hip::Role::Bulk) {
      reader = gpu::hip::parameterReader<T,DESCRIPTOR,DYNAMICS>(_parameters);
    }
    _state->add(typename gpu::hip::BlockState<T,DESCRIPTOR>::DynamicsEntry{
      info, _dynamics.get(), std::move(reader), _dynamics->getName()});
  }

  void apply(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>& block,
             ConcreteBlockMask<T,Platform::GPU_HIP>&,
             CollisionDispatchStrategy) override
  {
    _state->collide(block.statisticsEnabled());
  }

};

/// Cell-wise post-processors: those the device step runs are recorded, any
/// other one cannot run on GPU_HIP
template <typename T, typename DESCRIPTOR, concepts::CellOperator OPERATOR>
class ConcreteBlockO<T,DESCRIPTOR,Platform::GPU_HIP,OPERATOR,OperatorScope::PerCell> final
  : public BlockO<T,DESCRIPTOR,Platform::GPU_HIP> {
private:
  std::vector<CellID> _cells;
  std::shared_ptr<gpu::hip::BlockState<T,DESCRIPTOR>> _state;

public:
  std::type_index id() const override
  {
    return typeid(OPERATOR);
  }

  std::size_t weight() const override
  {
    return _cells.size();
  }

  void set(CellID iCell, bool state) override
  {
    if (state) {
      _cells.emplace_back(iCell);
      if (_state) {
        _state->markStructure();
      }
    }
  }

  void setup(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>& block) override
  {
    _state = gpu::hip::stateOf(block);
    constexpr gpu::hip::PostInfo info = gpu::hip::PostOf<OPERATOR>::info();
    if constexpr (info.handled) {
      _state->add(typename gpu::hip::BlockState<T,DESCRIPTOR>::PostEntry{info, &_cells, typeid(OPERATOR).name()});
    }
  }

  void apply(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>&) override
  {
    if constexpr (gpu::hip::PostOf<OPERATOR>::info().handled) {
      _state->postStream();
    } else {
      throw std::logic_error(std::string("GPU_HIP: post-processor ") + typeid(OPERATOR).name() +
                             " is not available on this platform");
    }
  }

};

template <typename T, typename DESCRIPTOR, concepts::CellOperator OPERATOR>
class ConcreteBlockO<T,DESCRIPTOR,Platform::GPU_HIP,OPERATOR,OperatorScope::PerCellWithParameters> final
  : public BlockO<T,DESCRIPTOR,Platform::GPU_HIP> {
private:
  std::vector<CellID> _cells;

public:
  std::type_index id() const override
  {
    return typeid(OPERATOR);
  }

  std::size_t weight() const override
  {
    return _cells.size();
  }

  void set(CellID iCell, bool state) override
  {
    if (state) {
      _cells.emplace_back(iCell);
    }
  }

  void setup(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>&) override { }

  void apply(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>&) override
  {
    throw std::logic_error(std::string("GPU_HIP: post-processor ") + typeid(OPERATOR).name() +
                           " is not available on this platform");
  }

};

/// Block-wise operators, e.g. StatisticsPostProcessor, run on the host block
template <typename T, typename DESCRIPTOR, concepts::BlockOperator OPERATOR>
class ConcreteBlockO<T,DESCRIPTOR,Platform::GPU_HIP,OPERATOR,OperatorScope::PerBlock> final
  : public BlockO<T,DESCRIPTOR,Platform::GPU_HIP> {
public:
  std::type_index id() const override
  {
    return typeid(OPERATOR);
  }

  std::size_t weight() const override
  {
    return 0;
  }

  void set(CellID, bool) override
  {
    throw std::logic_error("BlockO::set not supported for OperatorScope::PerBlock");
  }

  void setup(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>& block) override
  {
    OPERATOR().setup(block);
  }

  void apply(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>& block) override
  {
    OPERATOR().apply(block);
  }

};

/// Lattice statistics of a GPU_HIP block: the FP64 reduction the device ran
/// before this step's collision (lbmh_statistics), published where OpenLB's
/// StatisticsPostProcessor publishes the statistics gathered during collision
/// (src/core/postProcessing.hh:384-392, latticeStatistics.hh:75-90)
template <typename T, typename DESCRIPTOR>
struct StatisticsPostProcessor::type<ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>> {
  void setup(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>&) { }

  void apply(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>& block)
  {
    if (!block.statisticsEnabled()) {
      return;
    }
    const auto s = gpu::hip::stateOf(block)->statistics();
    block.getStatistics().reset(T(s[1]), T(s[2]), T(s[3]), std::size_t(s[0]));
  }
};

}

#endif
