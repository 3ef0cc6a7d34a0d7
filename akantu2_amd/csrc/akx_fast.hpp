// akx_fast.hpp -- entry points of the fast element pass (akx_fast.hip).
#pragma once
#include "akx_args.hpp"
#include "akx_batch.hpp"
#include <hip/hip_runtime.h>

namespace akxf {
// copy dN_n/dxi_m at the 8 Hex8 Gauss points ([q][m][n]) to the device
hipError_t upload_hex8_dnds(const double *dnds);
// elements handled by one block of the fast kernel of this element type
// elements per workgroup of the FP32 / FP64-fast element passes (bpa: the
// FP32 block-partial passes)
int64_t elements_per_block(int etype, bool bpa = false);
// launch the element pass of one material (etype 0 = Tet4, 1 = Hex8)
hipError_t launch_elements(int etype, int law, bool fp32, const akx::ElemArgs &A,
                           hipStream_t stream);
// launch the FP32-form batched element pass of one (element type, law) group;
// block = threads per block (64, 128 or 256; Hex8: 8 lanes per element)
hipError_t launch_batch_elements_f32(int etype, int law, int block,
                                     const akx::BatchElemArgs &A, hipStream_t stream);
} // namespace akxf
