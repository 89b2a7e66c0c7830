// FP32-storage LBM kernels (compiled in lbm_fast.hip with FMA contraction).
#pragma once
#include "lbm_device.h"

namespace psx_lbm_fast {

// One brick of the fused step: its nodes are [nstart, nstart + ncount); the
// colour field of its halo box is computed for the hcount nodes
// hnode[hstart..] into LDS cells hcell[hstart..]; (x0, y0, z0) is the voxel
// at box cell (1, 1, 1), i.e. the brick's first corner.
struct Brick {
    int32_t nstart, ncount, hstart, hcount;
    int32_t x0, y0, z0, pad_;
};

// Pull + collide over all nodes; `in`/`out` are AoSoA (T, s) blocks and
// `inx`/`outx` the extension bytes of the colour words (null: the plain
// 25-bit words).  The pull table is src16 (int16 residuals, -32768 =
// bounce-back) with the per-plane bases pbase[3][nx] when src16 is not null,
// else srca [nl][20] (node-major, two pad entries) when not null, else src
// [18][nl] (-1 = bounce-back); etab, when not null, replaces them all:
// [nl][20], the AoSoA element index of every pull (bounce-back resolved),
// the node's voxel and its coordinates x | y << 10 | z << 20.
// true when the collide kernel needs 64-bit pull offsets (the population
// buffer of nl nodes holds 2^31 floats or more; PSX_LBM_WIDE=1 forces it)
bool wide_index(int64_t nl);
hipError_t collide(const psx_lbm_dev::LbmConst &c, const float *in, const uint8_t *inx, float *out, uint8_t *outx,
                   const int32_t *src, const int16_t *src16, const int32_t *pbase, const int32_t *srca,
                   const uint32_t *etab, const int32_t *vox, const double *phi, int xcd, hipStream_t st);
// Pulled densities -> colour field phi at every node's voxel.
// dtab, when not null, is the int16 table node-major: [nl][16] offsets of
// directions 1..16, then [nl] x {offset 17, offset 18, voxel (int32)}.
hipError_t density(const psx_lbm_dev::LbmConst &c, const float *in, const uint8_t *inx, const int32_t *src,
                   const int16_t *src16, const int32_t *pbase, const uint16_t *dtab, const int32_t *vox,
                   double *phi, int xcd, hipStream_t st);
// Fused pull + colour + collide step over bricks (k_step_f32): one launch
// per time step.  box = the LDS box (brick + halo) dimensions (x, y, z).
hipError_t step(const psx_lbm_dev::LbmConst &c, const float *in, float *out, const int32_t *src,
                const int32_t *vox, const Brick *bricks, int nbricks, const int32_t *hnode, const int16_t *hcell,
                const int (&box)[3], int xcd, hipStream_t st);
// Per-link coupling term of the FP32 state (k_link_force).
hipError_t link_force(const psx_lbm_dev::LbmConst &c, const float *pop, const psx_lbm_dev::CplLinks &L, hipStream_t st);
// check_state partial sums (k_diag) of the pulled state.
hipError_t diag(const psx_lbm_dev::LbmConst &c, const float *in, const uint8_t *inx, const int32_t *src,
                const int32_t *vox, const double *phi, double *partial, hipStream_t st);

// save_results fields (k_fields) of the pulled state: out [5][nl], phase [nl].
hipError_t fields(const psx_lbm_dev::LbmConst &c, const float *in, const uint8_t *inx, const int32_t *src,
                  const int32_t *vox, const double *phi, double *out, int32_t *phase, hipStream_t st);

}  // namespace psx_lbm_fast
