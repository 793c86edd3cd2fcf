/* Declarations of Platform::GPU_HIP that OpenLB's generic core calls before
 * the platform's definitions are read (src/core/platform/gpu/hip.h, included
 * at the end of olb.h once every dynamics and post-processor type is known):
 * the block stream of ConcreteBlockLattice::stream and the device
 * synchronisation of SuperLattice::executePostProcessors.
 */
#ifndef OLB_GPU_HIP_FORWARD_H
#define OLB_GPU_HIP_FORWARD_H

namespace olb {

template <typename T, typename DESCRIPTOR, Platform PLATFORM> class ConcreteBlockLattice;

namespace gpu::hip {

template <typename T, typename DESCRIPTOR>
void stream(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_HIP>& block);

inline void synchronize();

}

}

#endif
