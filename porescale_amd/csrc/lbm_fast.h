// FP32-storage LBM kernels (compiled in lbm_fast.hip with FMA contraction).
#pragma once
#include "lbm_device.h"

namespace psx_lbm_fast {

// Collide + push over all nodes: `in` and `out` hold T[19][nl] then R[19][nl].
hipError_t collide(const psx_lbm_dev::LbmConst &c, const float *in, float *out, const int32_t *dst,
                   const int32_t *vox, const double *phi, int xcd, hipStream_t st);
// Densities -> colour field phi at every node's voxel.
hipError_t density(const psx_lbm_dev::LbmConst &c, const float *in, const int32_t *vox, double *phi, int xcd,
                   hipStream_t st);

}  // namespace psx_lbm_fast
