/* Columns of a Platform::GPU_HIP block.
 *
 * A GPU_HIP block keeps every field on the host, in the layout OpenLB's CPU
 * platforms use, and its populations additionally in the HBM columns of one
 * liblbmh lattice (include/lbmh.h).  Plain columns are OpenLB's host columns.
 * The population CyclicColumn is this file's own: a host column with the PS
 * shift (src/core/platform/cpu/sisd/column.h:128-221 is the contract) that,
 * once its block has a device lattice, copies itself to the device on a switch
 * to ProcessingContext::Simulation and back on a switch to Evaluation, as the
 * CUDA platform's CyclicColumn does (src/core/platform/gpu/cuda/column.hh:
 * 157-187, 285-296).
 */
#ifndef OLB_GPU_HIP_COLUMN_H
#define OLB_GPU_HIP_COLUMN_H

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <type_traits>

#include "lbmh.h"

namespace olb {

namespace gpu::hip {

/// Throws the library's last error, as gpu::cuda::device::check does for the CUDA runtime
inline void check(int status, const char* what)
{
  if (status != LBMH_OK) {
    throw std::runtime_error(std::string("GPU_HIP: ") + what + ": " + lbmh_last_error());
  }
}

/// OpenLB value type as the library's host dtype
template <typename T>
constexpr int hostType()
{
  static_assert(std::is_same_v<T,double> || std::is_same_v<T,float>,
                "GPU_HIP blocks hold double or float populations");
  return std::is_same_v<T,double> ? LBMH_F64 : LBMH_F32;
}

/// Receiver of a population column's context switches: the block's device lattice
struct PopulationMirror {
  virtual ~PopulationMirror() = default;
  /// host -> device: column iPop's values in storage order and its shift
  virtual void upload(unsigned iPop, const void* data, std::ptrdiff_t shift) = 0;
  /// device -> host, into the column's values at its current shift
  virtual void download(unsigned iPop, void* data, std::ptrdiff_t shift) = 0;
};

/// Population column of a GPU_HIP block: host storage with the PS shift,
/// mirrored by the block's device lattice while the block simulates
template <typename T>
class CyclicColumn final : public AbstractCyclicColumn<T>
                         , public Serializable {
private:
  std::size_t _count;
  std::unique_ptr<T[]> _data;
  std::ptrdiff_t _shift;

  PopulationMirror* _mirror;
  unsigned _iPop;

  std::size_t physical(std::size_t i) const
  {
    std::ptrdiff_t k = std::ptrdiff_t(i) + _shift;
    const std::ptrdiff_t n = std::ptrdiff_t(_count);
    if (k >= n) {
      k -= n;
    } else if (k < 0) {
      k += n;
    }
    return std::size_t(k);
  }

public:
  using value_t = T;

  CyclicColumn(std::size_t count):
    _count(count),
    _data(new T[count] {}),
    _shift(0),
    _mirror(nullptr),
    _iPop(0)
  { }

  CyclicColumn(CyclicColumn<T>&& rhs):
    _count(rhs._count),
    _data(std::move(rhs._data)),
    _shift(rhs._shift),
    _mirror(rhs._mirror),
    _iPop(rhs._iPop)
  {
    rhs._mirror = nullptr;
  }

  const T& operator[](std::size_t i) const override
  {
    return _data[physical(i)];
  }

  T& operator[](std::size_t i) override
  {
    return _data[physical(i)];
  }

  std::size_t size() const override
  {
    return _count;
  }

  T* data()
  {
    return _data.get();
  }

  const T* data() const
  {
    return _data.get();
  }

  std::ptrdiff_t getShift() const
  {
    return _shift;
  }

  /// Host-side propagation (src/core/platform/cpu/sisd/column.h:193-204);
  /// a simulating block streams on the device instead
  void rotate(std::ptrdiff_t offset)
  {
    const std::ptrdiff_t n = std::ptrdiff_t(_count);
    _shift -= offset;
    if (_shift >= n) {
      _shift -= n;
    } else if (_shift <= -n) {
      _shift += n;
    }
  }

  void refresh() { }

  void resize(std::size_t)
  {
    throw std::logic_error("Cyclic column can not be resized");
  }

  void attach(PopulationMirror* mirror, unsigned iPop)
  {
    _mirror = mirror;
    _iPop = iPop;
  }

  void setProcessingContext(ProcessingContext context)
  {
    if (!_mirror) {
      return;
    }
    if (context == ProcessingContext::Simulation) {
      _mirror->upload(_iPop, _data.get(), _shift);
    } else {
      _mirror->download(_iPop, _data.get(), _shift);
    }
  }

  /// Serialized as {shift, count, values}, the bytes of OpenLB's CPU
  /// CyclicColumn (src/core/platform/cpu/sisd/column.hh:66-89)
  std::size_t getNblock() const override
  {
    return 3;
  }

  std::size_t getSerializableSize() const override
  {
    return _count * sizeof(T) + sizeof(std::ptrdiff_t) + sizeof(std::size_t);
  }

  bool* getBlock(std::size_t iBlock, std::size_t& sizeBlock, bool loadingMode) override
  {
    std::size_t currentBlock = 0;
    bool* dataPtr = nullptr;
    registerVar(iBlock, sizeBlock, currentBlock, dataPtr, _shift);
    registerVar(iBlock, sizeBlock, currentBlock, dataPtr, _count);
    registerVar(iBlock, sizeBlock, currentBlock, dataPtr, *_data.get(), _count);
    return dataPtr;
  }

  void postLoad() override { }

};

}

/// Fields other than populations stay in OpenLB's host columns
template <typename T>
struct ImplementationOf<AbstractColumn<T>,Platform::GPU_HIP> {
  using type = cpu::sisd::Column<T>;
};

template <typename T>
struct ImplementationOf<AbstractCyclicColumn<T>,Platform::GPU_HIP> {
  using type = gpu::hip::CyclicColumn<T>;
};

}

#endif
