// Geometric multigrid preconditioner for the batched LSM solver.
//
// Level 0 is the batch's own node grid with the matrix-free element
// operator.  Level l+1 halves every axis of level l (ceil(n/2) coarse nodes
// on the even fine indices of the periodic axis; odd fine nodes interpolate
// from their two even neighbours), P is that trilinear interpolation and the
// coarse operators are the Galerkin products A_{l+1} = P^T A_l P, stored as
// 27-point stencils of 3x3 blocks.  Smoothing is Chebyshev on the
// block-Jacobi scaled operator D^-1 A over [lam/ratio, lam], lam per
// volume and level the smaller of the Gershgorin bound and 1.1 times a
// power-iteration estimate of the largest eigenvalue.  The V-cycle is symmetric (same
// polynomial before and after the coarse correction), so it is a valid PCG
// preconditioner.
#pragma once
#include "lsm_common.h"

namespace lsmk {

struct MGLevel {
    std::vector<VolDesc> vols;
    std::vector<int32_t> tvol, tstart;
    VolDesc *d_vols = nullptr;
    int32_t *d_tvol = nullptr, *d_tstart = nullptr;
    int ntiles = 0;
    int64_t N = 0;
    float *A = nullptr;     // [27*9][N] stencil (levels >= 1)
    // the same stencil for the smoothing steps: each node's 243 entries in
    // FP16 relative to the node's largest magnitude As[n] (null: use A)
    _Float16 *Ah = nullptr;
    float *As = nullptr;
    float *minv = nullptr;  // [6][N] inverse diagonal blocks
    float *lam = nullptr;   // [nvol] spectral bound of D^-1 A
    // work vectors [R][3][N]; level 0 uses the caller's r (b) and z (x)
    float *b = nullptr, *x = nullptr, *r = nullptr, *dA = nullptr, *dB = nullptr;
    bool own_bx = false, own_minv = false;
};

struct MGParams {
    int levels = 6;        // at most this many levels (including level 0)
    int coarse_max = 16;   // coarsen until the largest axis of every volume is at most this
    int nu = 3;            // Chebyshev degree of each smoothing pass (levels >= 1)
    int nu0 = 1;           // the same on the finest level (its matvec is the costliest)
    int coarse_deg = 4;    // Chebyshev degree on the coarsest level
    float ratio = 30.f;    // smoothing interval [lam/ratio, lam]
    float coarse_ratio = 200.f;
    // power-iteration steps per level for the spectral bound: 20 and 30 give
    // the same iteration counts at 64^3, 128^3 and on the uneven batch
    // (12 is 26-32 % slower at 128^3)
    int power_steps = 20;
    float power_safety = 1.1f;   // bound = min(Gershgorin, safety * power estimate)
    int rall0 = 1;               // finest-level smoothing with all right-hand sides per thread
};

struct MG {
    std::vector<MGLevel> lv;
    MGParams prm;
    int R = 0;
    bool built = false;
    int gen = 0;  // changes with every build (new buffers)
};

// Builds the hierarchy for the batch (R right-hand sides per volume).
int mg_build(hipStream_t st, const Batch &B, int R, const MGParams &prm, MG &mg);
void mg_free(MG &mg);
// Recomputes the operators of a built hierarchy from the batch's current
// spring constants (B.d_ka, B.d_kb, B.d_minv): coarse stencils, inverse
// diagonal blocks and spectral bounds.  Levels, tiles and work vectors stay.
int mg_operators(hipStream_t st, const Batch &B, MG &mg);
// The PCG update that mg_apply can fold into its first kernel:
// X += alpha P, r -= alpha Q (alpha per system), partials of r.r into prr.
struct CGUpdate {
    float *X;
    const float *P, *Q;
    double *prr;
};

// z = M^-1 r for every active system; r, z are [R][3][Ntot] FP32 on the
// batch grid.  prz (may be null) receives the (tile, rhs) partials of r.z.
// With upd, r is first updated in place as described above.
void mg_apply(hipStream_t st, const Batch &B, MG &mg, const SysState *sys, const float *r, float *z, double *prz,
              const CGUpdate *upd = nullptr);
// The launch variants mg_apply selects by environment (PSX_MG_RALL_TILES,
// PSX_MG_UYZ), read at each call: a captured graph keeps those of its capture.
int mg_launch_rall_tiles();
int mg_launch_uyz();
// FP16 stencils for the smoothing steps of levels >= 1 (PSX_MG_HALF, default
// 1), read when the hierarchy's operators are built
int mg_half_stencils();
// Parameters from the environment (PSX_MG_LEVELS, PSX_MG_CMAX, PSX_MG_NU, PSX_MG_NU0,
// PSX_MG_CDEG, PSX_MG_RATIO, PSX_MG_CRATIO, PSX_MG_POWER, PSX_MG_RALL0); defaults
// otherwise.
MGParams mg_params_from_env();

}  // namespace lsmk
