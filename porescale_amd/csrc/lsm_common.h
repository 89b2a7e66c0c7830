// Shared pieces of the LSM solver (lsm.hip) and its multigrid
// preconditioner (lsm_mg.hip): batch geometry, per-system scalars, the
// 27-node gather and tile helpers.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

#include "psx_common.h"

namespace lsmk {

constexpr int TILE = 256;
constexpr int MAXR = 6;  // right-hand sides per volume (load cases)

struct VolDesc {
    int32_t nx, ny, nz, pad_;
    int64_t off;      // node / voxel offset of the volume
    int32_t tile0;    // first tile of the volume
    int32_t ntile;
};

struct Jump {
    double J[MAXR][3][3];  // [rhs][crossed axis][component]
};

// per-system scalars (system s = volume*R + rhs)
struct SysState {
    double rz, pq, alpha, beta, rr, bb, scale;
    double bref;  // squared norm of the rhs summed without cancellation (0 = unused):
                  // a residual below 1e-14 of it is rounding noise and counts as converged
    double mean[3];  // outer refinement: mean residual over the nodes with springs
    double tol2;     // inner solves: squared relative tolerance of this system (0 = the launch's)
    int32_t active, iters, conv, pad2_;
};
constexpr double NOISE2 = 1e-28;

__device__ __forceinline__ int wrapm(int t, int n, int &w) {
    w = t < 0 ? -1 : (t >= n ? 1 : 0);
    return t < 0 ? t + n : (t >= n ? t - n : t);
}

// Gathers the 27-node neighbourhood of node (x,y,z) of one right-hand side.
template <typename T, bool JUMP>
__device__ __forceinline__ void gather27(const VolDesc &vd, const T *__restrict__ u, int64_t Ntot, int x, int y,
                                         int z, const double (&J)[3][3], T (&out)[27][3]) {
#pragma unroll
    for (int dz = -1; dz <= 1; dz++) {
        int wz, zz = wrapm(z + dz, vd.nz, wz);
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
            int wy, yy = wrapm(y + dy, vd.ny, wy);
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                int wx, xx = wrapm(x + dx, vd.nx, wx);
                int64_t idx = vd.off + xx + (int64_t)vd.nx * (yy + (int64_t)vd.ny * zz);
                int s = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    T v = u[(int64_t)e * Ntot + idx];
                    if (JUMP) v += (T)(wx * J[0][e] + wy * J[1][e] + wz * J[2][e]);
                    out[s][e] = v;
                }
            }
        }
    }
}

// Scales of the 8 voxels around node (x,y,z): ka[p] for voxel node - CORNER[p].
template <typename T>
__device__ __forceinline__ bool gather_k(const VolDesc &vd, const T *__restrict__ ka, const T *__restrict__ kb,
                                         int x, int y, int z, T (&a)[8], T (&b)[8]) {
    bool any = false;
#pragma unroll
    for (int p = 0; p < 8; p++) {
        int w, xx = wrapm(x - (p & 1), vd.nx, w), yy = wrapm(y - ((p >> 1) & 1), vd.ny, w),
               zz = wrapm(z - (p >> 2), vd.nz, w);
        int64_t v = vd.off + xx + (int64_t)vd.nx * (yy + (int64_t)vd.ny * zz);
        a[p] = ka[v];
        b[p] = kb[v];
        any |= (a[p] != T(0)) || (b[p] != T(0));
    }
    return any;
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
    v = psx::wave_sum(v);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sh[k];
    return s;  // valid in thread 0
}

// Tile shapes.  VolDesc::pad_ == 0: a tile is 256 consecutive nodes (x
// fastest) and `start` its first node.  Otherwise pad_ = 1 | lbx << 4 |
// lby << 8 | lbz << 12 describes a brick of 2^lbx x 2^lby x 2^lbz = 256
// nodes and `start` is the brick index (x fastest) within the volume.
inline int tile_shape_code(int lbx, int lby, int lbz) { return 1 | lbx << 4 | lby << 8 | lbz << 12; }

inline int64_t tiles_for(const VolDesc &d, int shape) {
    if (!(shape & 1)) return ((int64_t)d.nx * d.ny * d.nz + 255) / 256;
    const int lbx = (shape >> 4) & 15, lby = (shape >> 8) & 15, lbz = (shape >> 12) & 15;
    const int64_t nbx = (d.nx + (1 << lbx) - 1) >> lbx, nby = (d.ny + (1 << lby) - 1) >> lby,
                  nbz = (d.nz + (1 << lbz) - 1) >> lbz;
    return nbx * nby * nbz;
}

// First tile of a workgroup's grid-stride loop over tiles.  The hardware
// deals workgroups to the 8 XCDs round robin; remapped, the workgroups of
// one XCD take a contiguous range of tiles, so the 27-point halo of a tile
// mostly lies in tiles whose data the same L2 holds.
#ifndef PSX_LSM_XCD
#define PSX_LSM_XCD 1
#endif
__device__ __forceinline__ int lsm_tile0() {
    return PSX_LSM_XCD ? (int)psx::xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
}

__device__ __forceinline__ void tile_node(const VolDesc &vd, int start, int &x, int &y, int &z, bool &ok) {
    if (vd.pad_ & 1) {
        const int s = vd.pad_;
        const int lbx = (s >> 4) & 15, lby = (s >> 8) & 15, lbz = (s >> 12) & 15;
        const int nbx = (vd.nx + (1 << lbx) - 1) >> lbx, nby = (vd.ny + (1 << lby) - 1) >> lby;
        const int bxi = start % nbx, r2 = start / nbx, byi = r2 % nby, bzi = r2 / nby;
        const int tid = (int)threadIdx.x;
        x = (bxi << lbx) + (tid & ((1 << lbx) - 1));
        y = (byi << lby) + ((tid >> lbx) & ((1 << lby) - 1));
        z = (bzi << lbz) + (tid >> (lbx + lby));
        ok = x < vd.nx && y < vd.ny && z < vd.nz;
        if (!ok) x = y = z = 0;
        return;
    }
    int n = start + (int)threadIdx.x;
    int64_t N = (int64_t)vd.nx * vd.ny * vd.nz;
    ok = n < N;
    if (!ok) n = 0;
    x = n % vd.nx;
    int t = n / vd.nx;
    y = t % vd.ny;
    z = t / vd.ny;
}


struct Batch {
    int shape = 0;  // tile shape of every volume (VolDesc::pad_)
    std::vector<VolDesc> vols;
    std::vector<int32_t> tvol, tstart;
    int64_t Ntot = 0;
    int ntiles = 0;
    VolDesc *d_vols = nullptr;
    int32_t *d_tvol = nullptr, *d_tstart = nullptr;
    double *d_ka = nullptr, *d_kb = nullptr;
    float *d_ka32 = nullptr, *d_kb32 = nullptr;
    double *d_minv = nullptr;
    float *d_minv32 = nullptr;
    uint8_t *d_comp = nullptr;   // component per voxel (0 pore, 1, 2)
    double *d_stiff = nullptr;   // stiffness multiplier per voxel, or null (all 1)
    std::vector<uint8_t> comp_h;
    std::vector<double> stiff_h;  // empty: all 1
    double stiff_at(int64_t i) const { return stiff_h.empty() ? 1.0 : stiff_h[(size_t)i]; }
};

// Rebuilds the tile lists of every volume for a tile shape (host side).
inline void make_tiles(Batch &B, int shape) {
    B.shape = shape;
    B.tvol.clear();
    B.tstart.clear();
    int tile = 0;
    for (size_t v = 0; v < B.vols.size(); v++) {
        VolDesc &d = B.vols[v];
        d.pad_ = shape;
        d.tile0 = tile;
        d.ntile = (int)tiles_for(d, shape);
        for (int k = 0; k < d.ntile; k++) {
            B.tvol.push_back((int32_t)v);
            B.tstart.push_back(shape & 1 ? k : k * 256);
        }
        tile += d.ntile;
    }
    B.ntiles = tile;
}

// Workgroups for a tile loop: at most 8 per CU, one per tile.
inline int grid_for(int ntiles) {
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        hipDeviceProp_t prop;
        ncu = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
                  ? prop.multiProcessorCount : 256;
    }
    return std::max(1, std::min(ntiles, ncu * 8));
}

template <typename T>
int dalloc(T **p, size_t n) {
    PSX_HIP(hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)));
    return PSX_OK;
}

#define RC(x)                 \
    do {                      \
        int rc_ = (x);        \
        if (rc_) return rc_;  \
    } while (0)

}  // namespace lsmk
