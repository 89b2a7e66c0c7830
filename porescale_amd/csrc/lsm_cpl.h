// Internal interface (not exported) between the coupled LBM->LSM model in
// lbm.hip and the lattice spring model in lsm.hip: the solid lattice of the
// coupled model is an LSM handle over the same voxels with no imposed
// strain, advanced by lsm_step (L1:1721-1903) with an external force.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct psx_lsm;

namespace psx_lsm_cpl {

// Node count of the periodic node grid (nx*ny*nz); vectors are [3][N].
int64_t grid_nodes(const psx_lsm *h);
// Zeroes the coupled state r, v, a (L1:993-995) and prepares the operator.
int reset(psx_lsm *h);
// One lsm_step on stream st: v += a dt/2 + af dt/2; r += v dt;
// a = F(r) rsi - damp v; v += a dt/2 + af dt/2 (af: device [3][N]).
int step(psx_lsm *h, const double *d_af, double damp, hipStream_t st);
// Device pointers of r, v, a ([3][N]); valid after reset().
void state(psx_lsm *h, double **r, double **v, double **a);

}  // namespace psx_lsm_cpl
