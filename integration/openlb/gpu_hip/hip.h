/* Platform::GPU_HIP: OpenLB blocks whose time step runs on an AMD Instinct
 * MI355X (gfx950) through the C ABI of liblbmh (include/lbmh.h).
 *
 * Included by src/core/blockLattice.h where the CUDA platform's headers are
 * (src/core/platform/gpu/cuda.h).  The host code is plain C++: every device
 * kernel lives in liblbmh.so, so OpenLB stays a g++ build linked against one
 * C library.
 */
#ifndef OLB_GPU_HIP_H
#define OLB_GPU_HIP_H

#include "hip/column.h"
#include "hip/mask.h"
#include "hip/roles.h"
#include "hip/block.h"
#include "hip/operator.h"
#include "hip/communicator.h"

#endif
