/* Cell masks of a Platform::GPU_HIP block: host-only.  The device lattice
 * carries one dynamics-id byte per cell instead of a mask per dynamics
 * (include/lbmh.h, lbmh_set_dynamics_ids).  Interface: ConcreteBlockMask
 * (src/core/operator.h:32, used by BlockDynamicsMap).
 */
#ifndef OLB_GPU_HIP_MASK_H
#define OLB_GPU_HIP_MASK_H

namespace olb {

template <typename T>
class ConcreteBlockMask<T,Platform::GPU_HIP> final : public Serializable {
private:
  cpu::sisd::Column<bool> _mask;
  std::size_t _weight;

public:
  ConcreteBlockMask(std::size_t size):
    _mask(size),
    _weight(0)
  { }

  AbstractColumn<bool>& asAbstract()
  {
    return _mask;
  }

  bool operator[](std::size_t i) const
  {
    return _mask[i];
  }

  void set(std::size_t i, bool active)
  {
    if (_mask[i] != active) {
      _weight = active ? _weight + 1 : _weight - 1;
      _mask[i] = active;
    }
  }

  std::size_t weight() const
  {
    return _weight;
  }

  void setProcessingContext(ProcessingContext) { }

  std::size_t getNblock() const override
  {
    return 1 + _mask.getNblock();
  }

  std::size_t getSerializableSize() const override
  {
    return sizeof(_weight) + _mask.getSerializableSize();
  }

  bool* getBlock(std::size_t iBlock, std::size_t& sizeBlock, bool loadingMode) override
  {
    std::size_t currentBlock = 0;
    bool* dataPtr = nullptr;
    registerVar(iBlock, sizeBlock, currentBlock, dataPtr, _weight);
    registerSerializableOfConstSize(iBlock, sizeBlock, currentBlock, dataPtr, _mask, loadingMode);
    return dataPtr;
  }

};

}

#endif
