// Geometric multigrid preconditioner for the batched LSM solver (see
// lsm_mg.h for the construction).  All levels of all volumes of a batch are
// processed by the same launches: work is dealt as 256-node tiles of every
// volume, as in the solver.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "lsm_element.h"
#include "lsm_mg.h"

namespace lsmk {
namespace {

constexpr int CENTER = 13;  // stencil slot of offset (0,0,0)
constexpr int NSTEN = 27 * 9;  // coefficients per node: 27 offsets x 3x3 block

// ------------------------------------------------------------ coarsening --
// Coarse node I of an axis of n fine nodes sits on fine node 2I; odd fine
// nodes interpolate from their two (periodic) even neighbours.
__device__ inline int axis_support(int I, int n, int (&fi)[3], double (&w)[3]) {
    int k = 0;
    fi[k] = 2 * I; w[k] = 1.0; k++;
    if (2 * I + 1 < n) { fi[k] = 2 * I + 1; w[k] = 0.5; k++; }
    if (I >= 1) { fi[k] = 2 * I - 1; w[k] = 0.5; k++; }
    else if ((n & 1) == 0) { fi[k] = n - 1; w[k] = 0.5; k++; }
    return k;
}

__device__ inline int axis_parents(int j, int n, int (&J)[2], double (&w)[2]) {
    if ((j & 1) == 0) { J[0] = j >> 1; w[0] = 1.0; return 1; }
    J[0] = (j - 1) >> 1; w[0] = 0.5;
    int jp = j + 1 == n ? 0 : j + 1;
    J[1] = jp >> 1; w[1] = 0.5;
    return 2;
}

// stencil slot (-1, 0, 1) of coarse node J seen from I on a periodic axis of m >= 3 nodes
__device__ inline int cslot(int J, int I, int m) {
    int d = J - I;
    if (d > 1) d -= m;
    else if (d < -1) d += m;
    return d;
}

struct AxisCombo {
    int n;
    int i[12], d[12];
    double w[12];
};

// (fine node, fine offset, weight) triples of P^T A P coupling coarse I to
// the coarse neighbour in slot `target` along one axis.
__device__ inline void axis_combos(int I, int n, int m, int target, AxisCombo &c) {
    int fi[3];
    double fw[3];
    int ns = axis_support(I, n, fi, fw);
    c.n = 0;
    for (int a = 0; a < ns; a++)
        for (int dl = -1; dl <= 1; dl++) {
            int j = fi[a] + dl;
            j = j < 0 ? j + n : (j >= n ? j - n : j);
            int J[2];
            double wj[2];
            int np = axis_parents(j, n, J, wj);
            for (int b = 0; b < np; b++)
                if (cslot(J[b], I, m) == target && c.n < 12) {
                    c.i[c.n] = fi[a]; c.d[c.n] = dl; c.w[c.n] = fw[a] * wj[b]; c.n++;
                }
        }
}

__device__ inline int64_t node_index(const VolDesc &vd, int x, int y, int z) {
    return vd.off + x + (int64_t)vd.nx * (y + (int64_t)vd.ny * z);
}

__device__ inline int wrap1(int t, int n) { return t < 0 ? t + n : (t >= n ? t - n : t); }

// ------------------------------------------------------------- set-up -----
// Assembled fine operator: A0[(9s + 3e + e2) * N + n] = K(n, n + off(s))[e][e2].
__global__ __launch_bounds__(256) void k_assemble_fine(const VolDesc *__restrict__ vols,
                                                       const int32_t *__restrict__ tvol,
                                                       const int32_t *__restrict__ tstart, int ntiles,
                                                       const double *__restrict__ ka, const double *__restrict__ kb,
                                                       double *__restrict__ A0, int64_t N) {
    for (int t = lsm_tile0(); t < ntiles; t += gridDim.x) {
        const VolDesc vd = vols[tvol[t]];
        int x, y, z;
        bool ok;
        tile_node(vd, tstart[t], x, y, z, ok);
        if (!ok) continue;
        double a[8], b[8];
        gather_k(vd, ka, kb, x, y, z, a, b);
        lsm_el::node_stencil<double>(a, b, A0 + node_index(vd, x, y, z), N);
    }
}

// Galerkin product at coarse node (X, Y, Z) = I:
// Ac(I, slot) = sum P(i,I) Af(i, d) P(i+d, J).
__device__ void galerkin_node(const VolDesc &fd, const VolDesc &cd, int X, int Y, int Z, int64_t I,
                              const double *__restrict__ Af, int64_t Nf, double *__restrict__ Ac, int64_t Nc) {
        for (int sz = -1; sz <= 1; sz++) {
            AxisCombo cz;
            axis_combos(Z, fd.nz, cd.nz, sz, cz);
            for (int sy = -1; sy <= 1; sy++) {
                AxisCombo cy;
                axis_combos(Y, fd.ny, cd.ny, sy, cy);
                for (int sx = -1; sx <= 1; sx++) {
                    AxisCombo cx;
                    axis_combos(X, fd.nx, cd.nx, sx, cx);
                    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                    for (int kz = 0; kz < cz.n; kz++)
                        for (int ky = 0; ky < cy.n; ky++) {
                            double wyz = cz.w[kz] * cy.w[ky];
                            for (int kx = 0; kx < cx.n; kx++) {
                                double w = wyz * cx.w[kx];
                                int64_t fi = node_index(fd, cx.i[kx], cy.i[ky], cz.i[kz]);
                                int s = (cx.d[kx] + 1) + 3 * (cy.d[ky] + 1) + 9 * (cz.d[kz] + 1);
#pragma unroll
                                for (int k = 0; k < 9; k++) acc[k] += w * Af[(int64_t)(9 * s + k) * Nf + fi];
                            }
                        }
                    int sc = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
#pragma unroll
                    for (int k = 0; k < 9; k++) Ac[(int64_t)(9 * sc + k) * Nc + I] = acc[k];
                }
            }
        }
}

// Galerkin product on one level, from the assembled fine stencil.
__global__ __launch_bounds__(256) void k_galerkin(const VolDesc *__restrict__ fvols, const VolDesc *__restrict__ cvols,
                                                  const int32_t *__restrict__ ctvol,
                                                  const int32_t *__restrict__ ctstart, int cntiles,
                                                  const double *__restrict__ Af, int64_t Nf, double *__restrict__ Ac,
                                                  int64_t Nc) {
    for (int t = lsm_tile0(); t < cntiles; t += gridDim.x) {
        const int v = ctvol[t];
        const VolDesc fd = fvols[v], cd = cvols[v];
        int X, Y, Z;
        bool ok;
        tile_node(cd, ctstart[t], X, Y, Z, ok);
        if (!ok) continue;
        galerkin_node(fd, cd, X, Y, Z, node_index(cd, X, Y, Z), Af, Nf, Ac, Nc);
    }
}

// The first coarse level's Galerkin product from the element matrices.
// The fine operator is a sum of voxel (element) matrices a_v KA + b_v KB,
// and trilinear interpolation maps fine element i of an axis onto coarse
// element floor(i/2), its two fine corners taking the coarse corners'
// values with the weights [[1, 0], [1/2, 1/2]] (i even), [[1/2, 1/2],
// [0, 1]] (i odd), or [[1, 0], [0, 1]] (the last element of an odd-length
// periodic axis, the only fine element of the coarse element that wraps).
// So P^T A P = sum over fine voxels of a_v MA[p] + b_v MB[p] placed on the
// coarse element, p the voxel's pattern (one of 3 per axis) and M = P_p^T K
// P_p (mab: [27 p][2][24][24], exact dyadic values).  The coarse node sums,
// per stencil slot, the rows of its corner over the coarse elements that
// hold both ends: about 9k multiply-adds a node instead of the ~12k
// stencil loads of the general product, and no assembled fine stencil.
__global__ __launch_bounds__(256) void k_galerkin_elem(const VolDesc *__restrict__ fvols,
                                                       const VolDesc *__restrict__ cvols,
                                                       const int32_t *__restrict__ ctvol,
                                                       const int32_t *__restrict__ ctstart, int cntiles,
                                                       const double *__restrict__ ka, const double *__restrict__ kb,
                                                       const double *__restrict__ mab, double *__restrict__ Ac,
                                                       int64_t Nc) {
    for (int t = lsm_tile0(); t < cntiles; t += gridDim.x) {
        const int v = ctvol[t];
        const VolDesc fd = fvols[v], cd = cvols[v];
        int X, Y, Z;
        bool ok;
        tile_node(cd, ctstart[t], X, Y, Z, ok);
        if (!ok) continue;
        const int64_t I = node_index(cd, X, Y, Z);
        for (int sz = -1; sz <= 1; sz++)
            for (int sy = -1; sy <= 1; sy++)
                for (int sx = -1; sx <= 1; sx++) {
                    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                    // I sits at corner (cx, cy, cz) of a coarse element that
                    // also holds I + (sx, sy, sz): c = 1 for -1, 0 for +1,
                    // either for 0
                    for (int cz = sz == -1 ? 1 : 0; cz <= (sz == 1 ? 0 : 1); cz++)
                        for (int cy = sy == -1 ? 1 : 0; cy <= (sy == 1 ? 0 : 1); cy++)
                            for (int cx = sx == -1 ? 1 : 0; cx <= (sx == 1 ? 0 : 1); cx++) {
                                const int ex = wrap1(X - cx, cd.nx), ey = wrap1(Y - cy, cd.ny),
                                          ez = wrap1(Z - cz, cd.nz);
                                // the wrapping element of an odd fine axis holds one fine element
                                const bool wx = (fd.nx & 1) && ex == cd.nx - 1, wy = (fd.ny & 1) && ey == cd.ny - 1,
                                           wz = (fd.nz & 1) && ez == cd.nz - 1;
                                const int ci = cx + 2 * cy + 4 * cz;
                                const int cj = (cx + sx) + 2 * (cy + sy) + 4 * (cz + sz);
                                for (int kz = 0; kz < (wz ? 1 : 2); kz++)
                                    for (int ky = 0; ky < (wy ? 1 : 2); ky++)
                                        for (int kx = 0; kx < (wx ? 1 : 2); kx++) {
                                            const int64_t fv = node_index(fd, 2 * ex + kx, 2 * ey + ky, 2 * ez + kz);
                                            const double a = ka[fv], b = kb[fv];
                                            const int p = (wx ? 2 : kx) + 3 * (wy ? 2 : ky) + 9 * (wz ? 2 : kz);
                                            const double *ma = mab + ((size_t)(2 * p) * 24 + 3 * ci) * 24 + 3 * cj;
                                            const double *mb = ma + 24 * 24;
#pragma unroll
                                            for (int e = 0; e < 3; e++)
#pragma unroll
                                                for (int e2 = 0; e2 < 3; e2++)
                                                    acc[3 * e + e2] += a * ma[24 * e + e2] + b * mb[24 * e + e2];
                                        }
                            }
                    const int sc = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
#pragma unroll
                    for (int k = 0; k < 9; k++) Ac[(int64_t)(9 * sc + k) * Nc + I] = acc[k];
                }
    }
}

// mab of k_galerkin_elem: for each voxel pattern p = px + 3 py + 9 pz (per
// axis 0: even, 1: odd, 2: the wrapping element of an odd axis) the
// projected element matrices P_p^T KA P_p and P_p^T KB P_p, rows and
// columns 3 corner + component, corners x-bit first.
std::vector<double> galerkin_element_tables() {
    std::vector<double> m((size_t)27 * 2 * 24 * 24, 0.0);
    // [pattern][fine corner bit][coarse corner bit]
    const double P1[3][2][2] = {{{1.0, 0.0}, {0.5, 0.5}}, {{0.5, 0.5}, {0.0, 1.0}}, {{1.0, 0.0}, {0.0, 1.0}}};
    for (int p = 0; p < 27; p++) {
        const int pd[3] = {p % 3, (p / 3) % 3, p / 9};
        double P[8][8];
        for (int f = 0; f < 8; f++)
            for (int C = 0; C < 8; C++) {
                double w = 1.0;
                for (int d = 0; d < 3; d++) w *= P1[pd[d]][(f >> d) & 1][(C >> d) & 1];
                P[f][C] = w;
            }
        for (int k = 0; k < 2; k++) {
            const auto &K = k == 0 ? lsm_el::KA : lsm_el::KB;
            double *M = &m[((size_t)(2 * p + k)) * 24 * 24];
            for (int C = 0; C < 8; C++)
                for (int D = 0; D < 8; D++)
                    for (int e = 0; e < 3; e++)
                        for (int e2 = 0; e2 < 3; e2++) {
                            double acc = 0.0;
                            for (int f = 0; f < 8; f++)
                                for (int g = 0; g < 8; g++)
                                    acc += P[f][C] * K[3 * f + e][3 * g + e2] * P[g][D];
                            M[(3 * C + e) * 24 + 3 * D + e2] = acc;
                        }
        }
    }
    return m;
}

// FP16 copy of a level's stencil for the smoothing steps: node n's 243
// entries divided by their largest magnitude As[n] (1 for an all-zero
// row), so every entry keeps FP16's 11-bit significand whatever the
// springs' scale.  Halves the bytes the stencil steps stream; the power
// iteration and the Galerkin products keep the FP32 / FP64 stencils.
__global__ __launch_bounds__(256) void k_half_stencil(const float *__restrict__ A, int64_t N,
                                                      _Float16 *__restrict__ Ah, float *__restrict__ As) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float m = 0.f;
    for (int k = 0; k < NSTEN; k++) m = fmaxf(m, fabsf(A[(int64_t)k * N + n]));
    const float sc = m > 0.f ? m : 1.f;
    const float inv = 1.f / sc;
    for (int k = 0; k < NSTEN; k++) Ah[(int64_t)k * N + n] = (_Float16)(A[(int64_t)k * N + n] * inv);
    As[n] = sc;
}

__device__ inline void inv_block(const double (&D)[3][3], double (&m)[6]) {
    double c00 = D[1][1] * D[2][2] - D[1][2] * D[2][1];
    double c01 = D[1][2] * D[2][0] - D[1][0] * D[2][2];
    double c02 = D[1][0] * D[2][1] - D[1][1] * D[2][0];
    double det = D[0][0] * c00 + D[0][1] * c01 + D[0][2] * c02;
    for (int k = 0; k < 6; k++) m[k] = 0.0;
    double scale = fabs(D[0][0]) + fabs(D[1][1]) + fabs(D[2][2]);
    if (det > 1e-12 * scale * scale * scale && scale > 0.0) {
        double id = 1.0 / det;
        m[0] = c00 * id;
        m[1] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * id;
        m[2] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * id;
        m[3] = c01 * id;
        m[4] = c02 * id;
        m[5] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * id;
    } else {
        for (int e = 0; e < 3; e++) m[e] = D[e][e] > 0.0 ? 1.0 / D[e][e] : 0.0;
    }
}

// Per level: inverse diagonal blocks (unless given), FP32 copies and the
// Gershgorin bound max_i sum_j |(D^-1 A)_ij| as per-tile partial maxima.
__global__ __launch_bounds__(256) void k_level_prep(const VolDesc *__restrict__ vols, const int32_t *__restrict__ tvol,
                                                    const int32_t *__restrict__ tstart, int ntiles,
                                                    const double *__restrict__ A, int64_t N,
                                                    const double *__restrict__ minv_in, float *__restrict__ A32,
                                                    float *__restrict__ minv32, double *__restrict__ pmax) {
    __shared__ double sh[4];
    for (int t = lsm_tile0(); t < ntiles; t += gridDim.x) {
        const VolDesc vd = vols[tvol[t]];
        int x, y, z;
        bool ok;
        tile_node(vd, tstart[t], x, y, z, ok);
        double g = 0.0;
        if (ok) {
            const int64_t n = node_index(vd, x, y, z);
            double m[6];
            if (minv_in) {
                for (int k = 0; k < 6; k++) m[k] = minv_in[(int64_t)k * N + n];
            } else {
                double D[3][3];
                for (int e = 0; e < 3; e++)
                    for (int e2 = 0; e2 < 3; e2++) D[e][e2] = A[(int64_t)(9 * CENTER + 3 * e + e2) * N + n];
                inv_block(D, m);
            }
            if (minv32)
                for (int k = 0; k < 6; k++) minv32[(int64_t)k * N + n] = (float)m[k];
            const double M[3][3] = {{m[0], m[3], m[4]}, {m[3], m[1], m[5]}, {m[4], m[5], m[2]}};
            double rs[3] = {0, 0, 0};
            for (int s = 0; s < 27; s++) {
                double B[3][3];
                for (int e = 0; e < 3; e++)
                    for (int e2 = 0; e2 < 3; e2++) {
                        B[e][e2] = A[(int64_t)(9 * s + 3 * e + e2) * N + n];
                        if (A32) A32[(int64_t)(9 * s + 3 * e + e2) * N + n] = (float)B[e][e2];
                    }
                for (int e = 0; e < 3; e++)
                    for (int e2 = 0; e2 < 3; e2++)
                        rs[e] += fabs(M[e][0] * B[0][e2] + M[e][1] * B[1][e2] + M[e][2] * B[2][e2]);
            }
            g = fmax(rs[0], fmax(rs[1], rs[2]));
        }
        // block max
        for (int off = 32; off > 0; off >>= 1) g = fmax(g, __shfl_down(g, off, 64));
        int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        __syncthreads();
        if (lane == 0) sh[w] = g;
        __syncthreads();
        if (threadIdx.x == 0) {
            double mx = 0.0;
            for (int k = 0; k < (int)(blockDim.x >> 6); k++) mx = fmax(mx, sh[k]);
            pmax[t] = mx;
        }
    }
}

// Level 0's Gershgorin bound max_i sum_j |(D^-1 A)_ij| per tile, with the
// fine stencil's blocks formed from the voxel constants instead of read
// from an assembled stencil: block (n, n + s) = sum over the voxels holding
// both nodes of a_v KA + b_v KB at their corners (mab pattern 26, the
// identity interpolation, holds KA and KB themselves).
__global__ __launch_bounds__(256) void k_level_prep0_elem(const VolDesc *__restrict__ vols,
                                                          const int32_t *__restrict__ tvol,
                                                          const int32_t *__restrict__ tstart, int ntiles,
                                                          const double *__restrict__ ka,
                                                          const double *__restrict__ kb,
                                                          const double *__restrict__ mab, int64_t N,
                                                          const double *__restrict__ minv_in,
                                                          double *__restrict__ pmax) {
    __shared__ double sh[4];
    const double *KA = mab + (size_t)(2 * 26) * 24 * 24, *KB = KA + 24 * 24;
    for (int t = lsm_tile0(); t < ntiles; t += gridDim.x) {
        const VolDesc vd = vols[tvol[t]];
        int x, y, z;
        bool ok;
        tile_node(vd, tstart[t], x, y, z, ok);
        double g = 0.0;
        if (ok) {
            const int64_t n = node_index(vd, x, y, z);
            double m[6];
            for (int k = 0; k < 6; k++) m[k] = minv_in[(int64_t)k * N + n];
            const double M[3][3] = {{m[0], m[3], m[4]}, {m[3], m[1], m[5]}, {m[4], m[5], m[2]}};
            double rs[3] = {0, 0, 0};
            for (int sz = -1; sz <= 1; sz++)
                for (int sy = -1; sy <= 1; sy++)
                    for (int sx = -1; sx <= 1; sx++) {
                        double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                        for (int cz = sz == -1 ? 1 : 0; cz <= (sz == 1 ? 0 : 1); cz++)
                            for (int cy = sy == -1 ? 1 : 0; cy <= (sy == 1 ? 0 : 1); cy++)
                                for (int cx = sx == -1 ? 1 : 0; cx <= (sx == 1 ? 0 : 1); cx++) {
                                    const int64_t v = node_index(vd, wrap1(x - cx, vd.nx), wrap1(y - cy, vd.ny),
                                                                 wrap1(z - cz, vd.nz));
                                    const double a = ka[v], b = kb[v];
                                    const int ci = cx + 2 * cy + 4 * cz;
                                    const int cj = (cx + sx) + 2 * (cy + sy) + 4 * (cz + sz);
                                    for (int e = 0; e < 3; e++)
                                        for (int e2 = 0; e2 < 3; e2++)
                                            B[e][e2] += a * KA[(3 * ci + e) * 24 + 3 * cj + e2] +
                                                        b * KB[(3 * ci + e) * 24 + 3 * cj + e2];
                                }
                        for (int e = 0; e < 3; e++)
                            for (int e2 = 0; e2 < 3; e2++)
                                rs[e] += fabs(M[e][0] * B[0][e2] + M[e][1] * B[1][e2] + M[e][2] * B[2][e2]);
                    }
            g = fmax(rs[0], fmax(rs[1], rs[2]));
        }
        for (int off = 32; off > 0; off >>= 1) g = fmax(g, __shfl_down(g, off, 64));
        int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        __syncthreads();
        if (lane == 0) sh[w] = g;
        __syncthreads();
        if (threadIdx.x == 0) {
            double mx = 0.0;
            for (int k = 0; k < (int)(blockDim.x >> 6); k++) mx = fmax(mx, sh[k]);
            pmax[t] = mx;
        }
    }
}

__global__ __launch_bounds__(256) void k_vol_max(const VolDesc *__restrict__ vols, const double *__restrict__ pmax,
                                                 float *__restrict__ lam) {
    const VolDesc vd = vols[blockIdx.x];
    __shared__ double sh[4];
    double g = 0.0;
    for (int k = threadIdx.x; k < vd.ntile; k += blockDim.x) g = fmax(g, pmax[vd.tile0 + k]);
    for (int off = 32; off > 0; off >>= 1) g = fmax(g, __shfl_down(g, off, 64));
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = g;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) mx = fmax(mx, sh[k]);
        lam[blockIdx.x] = (float)mx;
    }
}

// --------------------------------------------------------- application ----
struct LevelArgs {
    const VolDesc *vols;
    const int32_t *tvol, *tstart;
    int ntiles;
    int64_t N;
    const float *A;     // stencil, or null for the element operator
    const _Float16 *Ah;  // the stencil in FP16 scaled per node by As (smoothing steps), or null
    const float *As;
    const float *ka, *kb;
    const float *minv;
    const float *lam;
    float ratio;
};

// Chebyshev coefficients of step k (k = -1: restart, d = D^-1 r / theta):
// d_{k+1} = c1 d_k + c2 D^-1 r_{k+1}.
__device__ inline void cheb_coef(float lam, float ratio, int k, float &c1, float &c2) {
    if (!(lam > 0.f)) { c1 = 0.f; c2 = 0.f; return; }
    const float beta = lam, alpha = lam / ratio;
    const float theta = 0.5f * (beta + alpha), delta = 0.5f * (beta - alpha);
    if (k < 0) { c1 = 0.f; c2 = 1.f / theta; return; }
    const float sigma = theta / delta;
    float rho = 1.f / sigma, rho_next = rho;
    for (int j = 0; j <= k; j++) {
        rho_next = 1.f / (2.f * sigma - rho);
        if (j < k) rho = rho_next;
    }
    c1 = rho_next * rho;
    c2 = 2.f * rho_next / delta;
}

__device__ inline void apply_minv(const float *__restrict__ minv, int64_t N, int64_t n, const float (&r)[3],
                                  float (&z)[3]) {
    float m0 = minv[n], m1 = minv[N + n], m2 = minv[2 * N + n], m3 = minv[3 * N + n], m4 = minv[4 * N + n],
          m5 = minv[5 * N + n];
    z[0] = m0 * r[0] + m3 * r[1] + m4 * r[2];
    z[1] = m3 * r[0] + m1 * r[1] + m5 * r[2];
    z[2] = m4 * r[0] + m5 * r[1] + m2 * r[2];
}

// (A u)(node) for one right-hand side u ([3][N], offset to that rhs).  UYZ 3
// unrolls the stencil's dy loop, putting nine neighbours' loads in flight
// (1.5-2% faster smoothing at 64^3; unrolling dz as well is 55% slower).
template <bool ELEM, int UYZ = 1, bool H = false>
__device__ inline void op_node(const LevelArgs &L, const VolDesc &vd, int x, int y, int z, int64_t n,
                               const float *__restrict__ u, bool act, const float (&a)[8], const float (&b)[8],
                               float (&out)[3]) {
    out[0] = out[1] = out[2] = 0.f;
    if (ELEM) {
        if (!act) return;
        float uu[27][3];
        const double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        gather27<float, false>(vd, u, L.N, x, y, z, J, uu);
        float f[3] = {0.f, 0.f, 0.f};
        lsm_el::node_force<float>(uu, a, b, f);
        out[0] = -f[0]; out[1] = -f[1]; out[2] = -f[2];
    } else {
#pragma unroll 1
        for (int dz = -1; dz <= 1; dz++) {
            int zz = wrap1(z + dz, vd.nz);
            constexpr int kUnrollY = UYZ >= 3 ? 3 : 1;
#pragma unroll kUnrollY
            for (int dy = -1; dy <= 1; dy++) {
                int yy = wrap1(y + dy, vd.ny);
#pragma unroll
                for (int dx = -1; dx <= 1; dx++) {
                    int xx = wrap1(x + dx, vd.nx);
                    int s = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                    int64_t nb = node_index(vd, xx, yy, zz);
                    float u0 = u[nb], u1 = u[L.N + nb], u2 = u[2 * L.N + nb];
#pragma unroll
                    for (int e = 0; e < 3; e++) {
                        if (H) {
                            const _Float16 *Ar = L.Ah + (int64_t)(9 * s + 3 * e) * L.N + n;
                            out[e] += (float)Ar[0] * u0 + (float)Ar[L.N] * u1 + (float)Ar[2 * L.N] * u2;
                        } else {
                            const float *Ar = L.A + (int64_t)(9 * s + 3 * e) * L.N + n;
                            out[e] += Ar[0] * u0 + Ar[L.N] * u1 + Ar[2 * L.N] * u2;
                        }
                    }
                }
            }
        }
        if (H) {
            const float sc = L.As[n];
            out[0] *= sc; out[1] *= sc; out[2] *= sc;
        }
    }
}

__device__ inline bool sys_active(const SysState *sys, int v, int R, int r) {
    return !sys || sys[v * R + r].active;
}

// FIRST: r = b, x = 0, d = D^-1 b / theta  (start of a smoothing pass from x = 0)
__global__ __launch_bounds__(TILE) void k_cheb_first(LevelArgs L, int R, const SysState *__restrict__ sys,
                                                     const float *__restrict__ bv, float *__restrict__ xv,
                                                     float *__restrict__ rv, float *__restrict__ dv) {
    for (int t = lsm_tile0(); t < L.ntiles; t += gridDim.x) {
        const int v = L.tvol[t];
        const VolDesc vd = L.vols[v];
        int x, y, z;
        bool ok;
        tile_node(vd, L.tstart[t], x, y, z, ok);
        if (!ok) continue;
        const int64_t n = node_index(vd, x, y, z);
        float c1, c2;
        cheb_coef(L.lam[v], L.ratio, -1, c1, c2);
        {  // one right-hand side per grid row
            const int r = blockIdx.y;
            if (!sys_active(sys, v, R, r)) continue;
            const int64_t base = (int64_t)r * 3 * L.N;
            float bb[3], zz[3];
            for (int e = 0; e < 3; e++) bb[e] = bv[base + e * L.N + n];
            apply_minv(L.minv, L.N, n, bb, zz);
            for (int e = 0; e < 3; e++) {
                if (xv) {  // no step follows: the pass is x = 0, r = b
                    xv[base + e * L.N + n] = 0.f;
                    rv[base + e * L.N + n] = bb[e];
                }
                dv[base + e * L.N + n] = c2 * zz[e];
            }
        }
    }
}

// The CG update fused with FIRST on level 0: x_cg += alpha p, r_cg -= alpha q
// (partials of r_cg.r_cg into prr), then r = r_cg, x = 0, d = D^-1 r_cg / theta.
__global__ __launch_bounds__(TILE) void k_cg_update_first(LevelArgs L, int R, const SysState *__restrict__ sys,
                                                          float *__restrict__ xcg, float *__restrict__ rcg,
                                                          const float *__restrict__ pcg, const float *__restrict__ qcg,
                                                          double *__restrict__ prr, float *__restrict__ xv,
                                                          float *__restrict__ rv, float *__restrict__ dv) {
    __shared__ double sh[TILE / 64];
    for (int t = lsm_tile0(); t < L.ntiles; t += gridDim.x) {
        const int v = L.tvol[t];
        const VolDesc vd = L.vols[v];
        int x, y, z;
        bool ok;
        tile_node(vd, L.tstart[t], x, y, z, ok);
        const int64_t n = node_index(vd, x, y, z);
        float c1, c2;
        cheb_coef(L.lam[v], L.ratio, -1, c1, c2);
        const int r = blockIdx.y;
        if (!sys_active(sys, v, R, r)) {
            if (threadIdx.x == 0) prr[(int64_t)t * R + r] = 0.0;
            continue;
        }
        const float al = (float)sys[v * R + r].alpha;
        const int64_t base = (int64_t)r * 3 * L.N;
        double drr = 0.0;
        if (ok) {
            float bb[3], zz[3];
            for (int e = 0; e < 3; e++) {
                const int64_t i = base + e * L.N + n;
                xcg[i] += al * pcg[i];
                float rn = rcg[i] - al * qcg[i];
                rcg[i] = rn;
                bb[e] = rn;
                drr += (double)rn * (double)rn;
            }
            apply_minv(L.minv, L.N, n, bb, zz);
            for (int e = 0; e < 3; e++) {
                const int64_t i = base + e * L.N + n;
                if (xv) {  // no step follows: the pass is x = 0, r = b
                    xv[i] = 0.f;
                    rv[i] = bb[e];
                }
                dv[i] = c2 * zz[e];
            }
        }
        double s = block_sum(drr, sh);
        if (threadIdx.x == 0) prr[(int64_t)t * R + r] = s;
    }
}

// STEP k: x += d (X0: x = d, the pass's first step from x = 0); r -= A d;
// (NEWD) d' = c1 d + c2 D^-1 r.  k = -1 restarts the recurrence
// (d' = D^-1 r / theta) after a coarse correction d.
template <bool ELEM, bool NEWD, int UYZ = 1, bool X0 = false, bool H = false>
__global__ __launch_bounds__(TILE) void k_cheb_step(LevelArgs L, int R, const SysState *__restrict__ sys, int k,
                                                    float *__restrict__ xv, float *__restrict__ rv,
                                                    const float *__restrict__ dv, float *__restrict__ dout,
                                                    const float *rin) {
    const float *rs = X0 && rin ? rin : rv;
    for (int t = lsm_tile0(); t < L.ntiles; t += gridDim.x) {
        const int v = L.tvol[t];
        const VolDesc vd = L.vols[v];
        int x, y, z;
        bool ok;
        tile_node(vd, L.tstart[t], x, y, z, ok);
        const int64_t n = node_index(vd, x, y, z);
        float a[8], b[8];
        bool act = false;
        if (ELEM && ok) act = gather_k(vd, L.ka, L.kb, x, y, z, a, b);
        float c1 = 0.f, c2 = 0.f;
        if (NEWD) cheb_coef(L.lam[v], L.ratio, k, c1, c2);
        {  // one right-hand side per grid row
            const int r = blockIdx.y;
            if (!sys_active(sys, v, R, r)) continue;
            const int64_t base = (int64_t)r * 3 * L.N;
            float q[3];
            op_node<ELEM, UYZ, H>(L, vd, x, y, z, n, dv + base, act, a, b, q);
            if (!ok) continue;
            // every load before the first store (rs may alias r)
            float rr[3], dd[3], xo[3], zz[3];
            for (int e = 0; e < 3; e++) {
                const int64_t i = base + e * L.N + n;
                dd[e] = dv[i];
                xo[e] = X0 ? 0.f : xv[i];
                rr[e] = rs[i] - q[e];
            }
            if (NEWD) apply_minv(L.minv, L.N, n, rr, zz);
            for (int e = 0; e < 3; e++) {
                const int64_t i = base + e * L.N + n;
                xv[i] = X0 ? dd[e] : xo[e] + dd[e];
                rv[i] = rr[e];
                if (NEWD) dout[i] = c1 * dd[e] + c2 * zz[e];
            }
        }
    }
}

// The same STEP on a stencil level for all right-hand sides at once: each
// thread reads its node's 27 stencil blocks (243 floats) once and applies
// them to every active right-hand side.  One grid row per right-hand side
// would stream the stencil R times per step; at 128^3 the first coarse
// level's stencil alone is 255 MB.  Reading only the 14 forward blocks and
// taking each backward block as the transpose of the neighbour's forward
// block (the Galerkin operator is symmetric) was measured 4-8 % slower: the
// neighbour's blocks miss the cache across tile edges.
template <bool NEWD, bool X0 = false, bool H = false>
__global__ __launch_bounds__(TILE) void k_cheb_step_rall(LevelArgs L, int R, const SysState *__restrict__ sys,
                                                         int k, float *__restrict__ xv, float *__restrict__ rv,
                                                         const float *__restrict__ dv, float *__restrict__ dout,
                                                         const float *rin) {
    const float *rs = X0 && rin ? rin : rv;
    for (int t = lsm_tile0(); t < L.ntiles; t += gridDim.x) {
        const int v = L.tvol[t];
        const VolDesc vd = L.vols[v];
        int x, y, z;
        bool ok;
        tile_node(vd, L.tstart[t], x, y, z, ok);
        if (!ok) continue;
        const int64_t n = node_index(vd, x, y, z);
        bool act[MAXR];
        bool any = false;
#pragma unroll
        for (int r = 0; r < MAXR; r++) {
            act[r] = r < R && sys_active(sys, v, R, r);
            any |= act[r];
        }
        if (!any) continue;
        float q[MAXR][3];
#pragma unroll
        for (int r = 0; r < MAXR; r++) q[r][0] = q[r][1] = q[r][2] = 0.f;
#pragma unroll 1
        for (int dz = -1; dz <= 1; dz++) {
            const int zz = wrap1(z + dz, vd.nz);
#pragma unroll 1
            for (int dy = -1; dy <= 1; dy++) {
                const int yy = wrap1(y + dy, vd.ny);
#pragma unroll
                for (int dx = -1; dx <= 1; dx++) {
                    const int xx = wrap1(x + dx, vd.nx);
                    const int s = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                    const int64_t nb = node_index(vd, xx, yy, zz);
                    float Ab[9];
#pragma unroll
                    for (int j = 0; j < 9; j++)
                        Ab[j] = H ? (float)L.Ah[(int64_t)(9 * s + j) * L.N + n] : L.A[(int64_t)(9 * s + j) * L.N + n];
                    // every right-hand side's loads unconditionally (an
                    // absent one, r >= R, reads right-hand side 0 and is
                    // dropped at the store), so the 18 loads of a
                    // neighbour are in flight together instead of three at
                    // a time behind a branch per right-hand side
#pragma unroll
                    for (int r = 0; r < MAXR; r++) {
                        const float *u = dv + (int64_t)(r < R ? r : 0) * 3 * L.N;
                        const float u0 = u[nb], u1 = u[L.N + nb], u2 = u[2 * L.N + nb];
#pragma unroll
                        for (int e = 0; e < 3; e++) q[r][e] += Ab[3 * e] * u0 + Ab[3 * e + 1] * u1 + Ab[3 * e + 2] * u2;
                    }
                }
            }
        }
        float c1 = 0.f, c2 = 0.f;
        if (NEWD) cheb_coef(L.lam[v], L.ratio, k, c1, c2);
        const float sc = H ? L.As[n] : 1.f;
        // every load of the update before its first store (r may alias rs,
        // which would otherwise hold each right-hand side's loads behind the
        // previous one's stores); absent right-hand sides read the first
        float dd[MAXR][3], xo[MAXR][3], rr[MAXR][3], mv[6];
#pragma unroll
        for (int r = 0; r < MAXR; r++) {
            const int64_t base = (int64_t)(r < R ? r : 0) * 3 * L.N;
#pragma unroll
            for (int e = 0; e < 3; e++) {
                const int64_t i = base + e * L.N + n;
                dd[r][e] = dv[i];
                xo[r][e] = X0 ? 0.f : xv[i];
                rr[r][e] = rs[i];
            }
        }
        if (NEWD)
#pragma unroll
            for (int j = 0; j < 6; j++) mv[j] = L.minv[j * L.N + n];
#pragma unroll
        for (int r = 0; r < MAXR; r++) {
            if (!act[r]) continue;
            const int64_t base = (int64_t)r * 3 * L.N;
            if (H) { q[r][0] *= sc; q[r][1] *= sc; q[r][2] *= sc; }
#pragma unroll
            for (int e = 0; e < 3; e++) {
                const int64_t i = base + e * L.N + n;
                xv[i] = X0 ? dd[r][e] : xo[r][e] + dd[r][e];
                rr[r][e] -= q[r][e];
                rv[i] = rr[r][e];
            }
            if (NEWD) {
                const float z0 = mv[0] * rr[r][0] + mv[3] * rr[r][1] + mv[4] * rr[r][2];
                const float z1 = mv[3] * rr[r][0] + mv[1] * rr[r][1] + mv[5] * rr[r][2];
                const float z2 = mv[4] * rr[r][0] + mv[5] * rr[r][1] + mv[2] * rr[r][2];
                dout[base + n] = c1 * dd[r][0] + c2 * z0;
                dout[base + L.N + n] = c1 * dd[r][1] + c2 * z1;
                dout[base + 2 * L.N + n] = c1 * dd[r][2] + c2 * z2;
            }
        }
    }
}

// Level-0 restart that also ends the V-cycle (smoothing degree 1 on the
// finest level): x += e; r -= A e; z = x + D^-1 r / theta, written to x (the
// caller's z), with the (tile, rhs) partials of r_cg.z into prz.
__global__ __launch_bounds__(TILE) void k_restart_final0(LevelArgs L, int R, const SysState *__restrict__ sys,
                                                         float *__restrict__ xv, float *__restrict__ rv,
                                                         const float *__restrict__ ev, const float *__restrict__ rcg,
                                                         double *__restrict__ prz) {
    __shared__ double sh[TILE / 64];
    for (int t = lsm_tile0(); t < L.ntiles; t += gridDim.x) {
        const int v = L.tvol[t];
        const VolDesc vd = L.vols[v];
        int x, y, z;
        bool ok;
        tile_node(vd, L.tstart[t], x, y, z, ok);
        const int64_t n = node_index(vd, x, y, z);
        const int r = blockIdx.y;
        if (!sys_active(sys, v, R, r)) {
            if (prz && threadIdx.x == 0) prz[(int64_t)t * R + r] = 0.0;
            continue;
        }
        float a[8], b[8];
        const bool act = ok && gather_k(vd, L.ka, L.kb, x, y, z, a, b);
        float c1, c2;
        cheb_coef(L.lam[v], L.ratio, -1, c1, c2);
        const int64_t base = (int64_t)r * 3 * L.N;
        float q[3];
        op_node<true>(L, vd, x, y, z, n, ev + base, act, a, b, q);
        double dot = 0.0;
        if (ok) {
            float rr[3], xx[3], zz[3];
            for (int e = 0; e < 3; e++) {
                const int64_t i = base + e * L.N + n;
                xx[e] = xv[i] + ev[i];
                rr[e] = rv[i] - q[e];
            }
            apply_minv(L.minv, L.N, n, rr, zz);
            for (int e = 0; e < 3; e++) {
                const int64_t i = base + e * L.N + n;
                const float zf = xx[e] + c2 * zz[e];
                xv[i] = zf;
                dot += (double)rcg[i] * (double)zf;
            }
        }
        if (prz) {
            double s = block_sum(dot, sh);
            if (threadIdx.x == 0) prz[(int64_t)t * R + r] = s;
        }
    }
}

// The finest level's two element-operator smoothing kernels with every
// right-hand side in one thread, the voxel scales gathered once per node
// (the CG matvec's layout), instead of one grid row per right-hand side:
//   FINAL = false: the first step of a degree-1 pass from x = 0:
//     x = d, r = b - A d (k_cheb_step<true, false, 1, true>);
//   FINAL = true: the restart that ends the V-cycle: z = x + e +
//     D^-1 (r - A e) / theta written to x, with the (tile, rhs) partials of
//     r_cg.z into prz (k_restart_final0).
template <bool FINAL>
__global__ __launch_bounds__(TILE) void k_step0_rall(LevelArgs L, int R, const SysState *__restrict__ sys,
                                                     float *__restrict__ xv, float *__restrict__ rv,
                                                     const float *__restrict__ dv, const float *__restrict__ rin,
                                                     const float *__restrict__ rcg, double *__restrict__ prz) {
    __shared__ double sh[TILE / 64];
    for (int t = lsm_tile0(); t < L.ntiles; t += gridDim.x) {
        const int v = L.tvol[t];
        const VolDesc vd = L.vols[v];
        int x, y, z;
        bool ok;
        tile_node(vd, L.tstart[t], x, y, z, ok);
        const int64_t n = node_index(vd, x, y, z);
        float a[8], b[8];
        const bool act = ok && gather_k(vd, L.ka, L.kb, x, y, z, a, b);
        float c1 = 0.f, c2 = 0.f;
        if (FINAL) cheb_coef(L.lam[v], L.ratio, -1, c1, c2);
        for (int r = 0; r < R; r++) {
            if (!sys_active(sys, v, R, r)) {  // uniform over the tile (one volume)
                if (FINAL && prz && threadIdx.x == 0) prz[(int64_t)t * R + r] = 0.0;
                continue;
            }
            const int64_t base = (int64_t)r * 3 * L.N;
            float q[3];
            op_node<true>(L, vd, x, y, z, n, dv + base, act, a, b, q);
            double dot = 0.0;
            if (ok) {
                if (!FINAL) {
                    for (int e = 0; e < 3; e++) {
                        const int64_t i = base + e * L.N + n;
                        xv[i] = dv[i];
                        rv[i] = rin[i] - q[e];
                    }
                } else {
                    float rr[3], xx[3], zz[3];
                    for (int e = 0; e < 3; e++) {
                        const int64_t i = base + e * L.N + n;
                        xx[e] = xv[i] + dv[i];
                        rr[e] = rv[i] - q[e];
                    }
                    apply_minv(L.minv, L.N, n, rr, zz);
                    for (int e = 0; e < 3; e++) {
                        const int64_t i = base + e * L.N + n;
                        const float zf = xx[e] + c2 * zz[e];
                        xv[i] = zf;
                        dot += (double)rcg[i] * (double)zf;
                    }
                }
            }
            if (FINAL && prz) {
                double s = block_sum(dot, sh);
                if (threadIdx.x == 0) prz[(int64_t)t * R + r] = s;
            }
        }
    }
}

// coarse node I: r = P^T r_fine (the level's right-hand side; no copy is
// kept); x = 0 when xc is given; d = D^-1 r / theta
__global__ __launch_bounds__(TILE) void k_restrict_first(LevelArgs F, LevelArgs C, int R,
                                                         const SysState *__restrict__ sys,
                                                         const float *__restrict__ rf, float *__restrict__ xc,
                                                         float *__restrict__ rc, float *__restrict__ dc) {
    for (int t = lsm_tile0(); t < C.ntiles; t += gridDim.x) {
        const int v = C.tvol[t];
        const VolDesc cd = C.vols[v], fd = F.vols[v];
        int X, Y, Z;
        bool ok;
        tile_node(cd, C.tstart[t], X, Y, Z, ok);
        if (!ok) continue;
        const int64_t I = node_index(cd, X, Y, Z);
        int ix[3], iy[3], iz[3];
        double wx[3], wy[3], wz[3];
        int nx = axis_support(X, fd.nx, ix, wx), ny = axis_support(Y, fd.ny, iy, wy),
            nz = axis_support(Z, fd.nz, iz, wz);
        float c1, c2;
        cheb_coef(C.lam[v], C.ratio, -1, c1, c2);
        {  // one right-hand side per grid row
            const int r = blockIdx.y;
            if (!sys_active(sys, v, R, r)) continue;
            const float *rr = rf + (int64_t)r * 3 * F.N;
            float s[3] = {0.f, 0.f, 0.f};
            for (int a = 0; a < nz; a++)
                for (int b = 0; b < ny; b++)
                    for (int c = 0; c < nx; c++) {
                        float w = (float)(wz[a] * wy[b] * wx[c]);
                        int64_t i = node_index(fd, ix[c], iy[b], iz[a]);
                        s[0] += w * rr[i];
                        s[1] += w * rr[F.N + i];
                        s[2] += w * rr[2 * F.N + i];
                    }
            float zz[3];
            apply_minv(C.minv, C.N, I, s, zz);
            const int64_t base = (int64_t)r * 3 * C.N;
            for (int e = 0; e < 3; e++) {
                rc[base + e * C.N + I] = s[e];
                if (xc) xc[base + e * C.N + I] = 0.f;
                dc[base + e * C.N + I] = c2 * zz[e];
            }
        }
    }
}

// fine node i: e = P (xc + dpend)
__global__ __launch_bounds__(TILE) void k_prolong(LevelArgs F, LevelArgs C, int R, const SysState *__restrict__ sys,
                                                  const float *__restrict__ xc, const float *__restrict__ dpend,
                                                  float *__restrict__ ef) {
    for (int t = lsm_tile0(); t < F.ntiles; t += gridDim.x) {
        const int v = F.tvol[t];
        const VolDesc fd = F.vols[v], cd = C.vols[v];
        int x, y, z;
        bool ok;
        tile_node(fd, F.tstart[t], x, y, z, ok);
        if (!ok) continue;
        const int64_t n = node_index(fd, x, y, z);
        int jx[2], jy[2], jz[2];
        double wx[2], wy[2], wz[2];
        int nx = axis_parents(x, fd.nx, jx, wx), ny = axis_parents(y, fd.ny, jy, wy),
            nz = axis_parents(z, fd.nz, jz, wz);
        {  // one right-hand side per grid row
            const int r = blockIdx.y;
            if (!sys_active(sys, v, R, r)) continue;
            const int64_t cb = (int64_t)r * 3 * C.N;
            float s[3] = {0.f, 0.f, 0.f};
            for (int a = 0; a < nz; a++)
                for (int b = 0; b < ny; b++)
                    for (int c = 0; c < nx; c++) {
                        float w = (float)(wz[a] * wy[b] * wx[c]);
                        int64_t J = node_index(cd, jx[c], jy[b], jz[a]);
                        for (int e = 0; e < 3; e++) {
                            float val = xc[cb + e * C.N + J];
                            if (dpend) val += dpend[cb + e * C.N + J];
                            s[e] += w * val;
                        }
                    }
            const int64_t fb = (int64_t)r * 3 * F.N;
            for (int e = 0; e < 3; e++) ef[fb + e * F.N + n] = s[e];
        }
    }
}

// level 0: z += d (the pending last update); partials of r.z per (tile, rhs)
__global__ __launch_bounds__(TILE) void k_final0(LevelArgs L, int R, const SysState *__restrict__ sys,
                                                 const float *__restrict__ rcg, float *__restrict__ z,
                                                 const float *__restrict__ d, double *__restrict__ prz) {
    __shared__ double sh[TILE / 64];
    for (int t = lsm_tile0(); t < L.ntiles; t += gridDim.x) {
        const int v = L.tvol[t];
        const VolDesc vd = L.vols[v];
        int x, y, zc;
        bool ok;
        tile_node(vd, L.tstart[t], x, y, zc, ok);
        const int64_t n = node_index(vd, x, y, zc);
        {  // one right-hand side per grid row
            const int r = blockIdx.y;
            if (!sys_active(sys, v, R, r)) {
                if (prz && threadIdx.x == 0) prz[(int64_t)t * R + r] = 0.0;
                continue;
            }
            double dot = 0.0;
            if (ok) {
                const int64_t base = (int64_t)r * 3 * L.N;
                for (int e = 0; e < 3; e++) {
                    const int64_t i = base + e * L.N + n;
                    float zv = z[i] + d[i];
                    z[i] = zv;
                    dot += (double)rcg[i] * (double)zv;
                }
            }
            if (prz) {
                double s = block_sum(dot, sh);
                if (threadIdx.x == 0) prz[(int64_t)t * R + r] = s;
            }
        }
    }
}

// ---------------------------------------------------- spectral bounds ---
// Power iteration on D^-1 A (right-hand side slot 0 of the level's work
// vectors): vout = D^-1 A vin / |vin|, per-tile partials of |vout|^2.
__global__ __launch_bounds__(TILE) void k_pow_init(LevelArgs L, float *__restrict__ v) {
    for (int t = lsm_tile0(); t < L.ntiles; t += gridDim.x) {
        const VolDesc vd = L.vols[L.tvol[t]];
        int x, y, z;
        bool ok;
        tile_node(vd, L.tstart[t], x, y, z, ok);
        if (!ok) continue;
        const int64_t n = node_index(vd, x, y, z);
        for (int e = 0; e < 3; e++) {
            uint32_t h = (uint32_t)(n * 3 + e) * 2654435761u;
            h ^= h >> 15; h *= 2246822519u; h ^= h >> 13;
            v[e * L.N + n] = (float)(h & 0xFFFF) / 65536.f - 0.5f;
        }
    }
}

template <bool ELEM>
__global__ __launch_bounds__(TILE) void k_pow_step(LevelArgs L, const float *__restrict__ vin, float *__restrict__ vout,
                                                   const double *__restrict__ nin, double *__restrict__ pnorm) {
    __shared__ double sh[TILE / 64];
    for (int t = lsm_tile0(); t < L.ntiles; t += gridDim.x) {
        const int v = L.tvol[t];
        const VolDesc vd = L.vols[v];
        int x, y, z;
        bool ok;
        tile_node(vd, L.tstart[t], x, y, z, ok);
        const int64_t n = node_index(vd, x, y, z);
        float a[8], b[8];
        bool act = false;
        if (ELEM && ok) act = gather_k(vd, L.ka, L.kb, x, y, z, a, b);
        float q[3];
        op_node<ELEM>(L, vd, x, y, z, n, vin, act, a, b, q);
        double d = 0.0;
        if (ok) {
            const float sc = nin && nin[v] > 0.0 ? (float)(1.0 / nin[v]) : 1.f;
            float qs[3] = {q[0] * sc, q[1] * sc, q[2] * sc}, zz[3];
            apply_minv(L.minv, L.N, n, qs, zz);
            for (int e = 0; e < 3; e++) {
                vout[e * L.N + n] = zz[e];
                d += (double)zz[e] * zz[e];
            }
        }
        double s = block_sum(d, sh);
        if (threadIdx.x == 0) pnorm[t] = s;
    }
}

// per volume: nrm = sqrt(sum of the tile partials)
__global__ __launch_bounds__(256) void k_vol_norm(const VolDesc *__restrict__ vols, const double *__restrict__ pnorm,
                                                  double *__restrict__ nrm) {
    const VolDesc vd = vols[blockIdx.x];
    __shared__ double sh[4];
    double g = 0.0;
    for (int k = threadIdx.x; k < vd.ntile; k += blockDim.x) g += pnorm[vd.tile0 + k];
    double s = block_sum(g, sh);
    if (threadIdx.x == 0) nrm[blockIdx.x] = sqrt(s);
}

// lam = min(Gershgorin bound, safety * power estimate)
__global__ void k_pow_lam(int nvol, const double *__restrict__ nrm, float *__restrict__ lam, float safety) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nvol && nrm[v] > 0.0) lam[v] = fminf(lam[v], safety * (float)nrm[v]);
}

LevelArgs args_of(const MGLevel &l, const Batch &B, float ratio) {
    LevelArgs a{};
    a.vols = l.d_vols; a.tvol = l.d_tvol; a.tstart = l.d_tstart; a.ntiles = l.ntiles; a.N = l.N;
    a.A = l.A; a.Ah = l.Ah; a.As = l.As; a.ka = B.d_ka32; a.kb = B.d_kb32; a.minv = l.minv; a.lam = l.lam; a.ratio = ratio;
    return a;
}

int env_int(const char *name, int def);

template <bool NEWD, bool X0>
void launch_step_t(hipStream_t st, const LevelArgs &L, bool elem, int R, const SysState *sys, int k, float *x,
                   float *r, const float *d, float *dout, const float *rin) {
    int g = grid_for(L.ntiles);
    // All right-hand sides per thread only where the level alone fills the
    // chip: with fewer tiles the R grid rows are the parallelism (64^3:
    // 2x slower with it on every level, 128^3: 15% faster on the 64^3 level).
    // read at every launch, so a process can compare the variants; a graph
    // records the variant of its capture (the solver's graph key holds both)
    const int rall_tiles = mg_launch_rall_tiles();
    const int uyz = mg_launch_uyz();
    const bool h = L.Ah != nullptr;
    if (elem)
        hipLaunchKernelGGL((k_cheb_step<true, NEWD, 1, X0>), dim3(g, R), dim3(TILE), 0, st, L, R, sys, k, x, r, d,
                           dout, rin);
    else if (rall_tiles > 0 && L.ntiles >= rall_tiles && h)
        hipLaunchKernelGGL((k_cheb_step_rall<NEWD, X0, true>), dim3(g), dim3(TILE), 0, st, L, R, sys, k, x, r, d,
                           dout, rin);
    else if (rall_tiles > 0 && L.ntiles >= rall_tiles)
        hipLaunchKernelGGL((k_cheb_step_rall<NEWD, X0>), dim3(g), dim3(TILE), 0, st, L, R, sys, k, x, r, d, dout,
                           rin);
    else if (uyz >= 3 && h)
        hipLaunchKernelGGL((k_cheb_step<false, NEWD, 3, X0, true>), dim3(g, R), dim3(TILE), 0, st, L, R, sys, k, x,
                           r, d, dout, rin);
    else if (uyz >= 3)
        hipLaunchKernelGGL((k_cheb_step<false, NEWD, 3, X0>), dim3(g, R), dim3(TILE), 0, st, L, R, sys, k, x, r, d,
                           dout, rin);
    else if (h)
        hipLaunchKernelGGL((k_cheb_step<false, NEWD, 1, X0, true>), dim3(g, R), dim3(TILE), 0, st, L, R, sys, k, x,
                           r, d, dout, rin);
    else
        hipLaunchKernelGGL((k_cheb_step<false, NEWD, 1, X0>), dim3(g, R), dim3(TILE), 0, st, L, R, sys, k, x, r, d,
                           dout, rin);
}

// x0: the level's x is zero (the first step of a smoothing pass from x = 0),
// so the step stores x = d without reading x; rin (first step only, else
// null): the pass's right-hand side, read in place of r, which the step
// then writes
template <bool NEWD>
void launch_step(hipStream_t st, const LevelArgs &L, bool elem, int R, const SysState *sys, int k, float *x, float *r,
                 const float *d, float *dout, bool x0, const float *rin = nullptr) {
    if (x0) launch_step_t<NEWD, true>(st, L, elem, R, sys, k, x, r, d, dout, rin);
    else launch_step_t<NEWD, false>(st, L, elem, R, sys, k, x, r, d, dout, nullptr);
}

// V-cycle from level l, whose b, r, dA hold a started smoothing pass
// (r = b, dA = d_0) and whose x is taken as 0: the first step stores x = d,
// so x needs zeroing only on a level that takes no step (a coarsest level of
// degree 1, see x_untouched).  Returns the pending last update of x.
// Level l takes no smoothing step in a V-cycle, so its x must start zeroed.
bool x_untouched(const MG &mg, int l) { return l + 1 == (int)mg.lv.size() && mg.prm.coarse_deg < 2; }

float *vcycle(hipStream_t st, const Batch &B, MG &mg, const SysState *sys, int l, const float *rcg = nullptr,
              double *prz = nullptr, bool *finished = nullptr) {
    MGLevel &lv = mg.lv[l];
    const bool coarsest = l + 1 == (int)mg.lv.size();
    const bool elem = l == 0;
    const MGParams &p = mg.prm;
    LevelArgs L = args_of(lv, B, coarsest ? p.coarse_ratio : p.ratio);
    float *dA = lv.dA, *dB = lv.dB;
    bool x0 = true;
    // level 0's pass starts from r = b without a copy: its first step reads b
    const float *rin = l == 0 ? lv.b : nullptr;
    if (coarsest) {
        for (int k = 0; k + 1 < p.coarse_deg; k++) {
            launch_step<true>(st, L, elem, mg.R, sys, k, lv.x, lv.r, dA, dB, x0, rin);
            x0 = false;
            std::swap(dA, dB);
        }
        return dA;
    }
    const int nu = l == 0 ? p.nu0 : p.nu;
    // pre-smoothing: nu updates, the last one with the residual update
    for (int k = 0; k + 1 < nu; k++) {
        launch_step<true>(st, L, elem, mg.R, sys, k, lv.x, lv.r, dA, dB, x0, rin);
        x0 = false;
        std::swap(dA, dB);
    }
    if (elem && nu == 1 && p.rall0)  // x = d, r = b - A d with all right-hand sides per thread
        hipLaunchKernelGGL(k_step0_rall<false>, dim3(grid_for(L.ntiles)), dim3(TILE), 0, st, L, mg.R, sys, lv.x,
                           lv.r, (const float *)dA, (const float *)rin, (const float *)nullptr, (double *)nullptr);
    else
        launch_step<false>(st, L, elem, mg.R, sys, 0, lv.x, lv.r, dA, nullptr, x0, x0 ? rin : nullptr);
    // coarse correction
    MGLevel &cv = mg.lv[l + 1];
    const bool c_coarsest = l + 2 == (int)mg.lv.size();
    LevelArgs C = args_of(cv, B, c_coarsest ? p.coarse_ratio : p.ratio);
    hipLaunchKernelGGL(k_restrict_first, dim3(grid_for(C.ntiles), mg.R), dim3(TILE), 0, st, L, C, mg.R, sys, lv.r,
                       x_untouched(mg, l + 1) ? cv.x : nullptr, cv.r, cv.dA);
    float *pend = vcycle(st, B, mg, sys, l + 1);
    hipLaunchKernelGGL(k_prolong, dim3(grid_for(L.ntiles), mg.R), dim3(TILE), 0, st, L, C, mg.R, sys, cv.x,
                       (const float *)pend, dA);
    // post-smoothing: restart from the corrected x, nu updates
    if (l == 0 && nu == 1 && finished) {  // the restart is the last update: it also forms z and r.z
        if (p.rall0)
            hipLaunchKernelGGL(k_step0_rall<true>, dim3(grid_for(L.ntiles)), dim3(TILE), 0, st, L, mg.R, sys, lv.x,
                               lv.r, (const float *)dA, (const float *)nullptr, rcg, prz);
        else
            hipLaunchKernelGGL(k_restart_final0, dim3(grid_for(L.ntiles), mg.R), dim3(TILE), 0, st, L, mg.R, sys,
                               lv.x, lv.r, (const float *)dA, rcg, prz);
        *finished = true;
        return nullptr;
    }
    launch_step<true>(st, L, elem, mg.R, sys, -1, lv.x, lv.r, dA, dB, false);
    std::swap(dA, dB);
    for (int k = 0; k + 1 < nu; k++) {
        launch_step<true>(st, L, elem, mg.R, sys, k, lv.x, lv.r, dA, dB, false);
        std::swap(dA, dB);
    }
    return dA;
}

int alloc_level_vectors(MGLevel &l, int R, bool bx) {
    size_t n = (size_t)R * 3 * l.N;
    if (bx) { RC(dalloc(&l.x, n)); l.own_bx = true; }  // a coarse level keeps no b (restriction writes r)
    RC(dalloc(&l.r, n)); RC(dalloc(&l.dA, n)); RC(dalloc(&l.dB, n));
    return PSX_OK;
}

int upload_level(hipStream_t st, MGLevel &l) {
    RC(dalloc(&l.d_vols, l.vols.size()));
    RC(dalloc(&l.d_tvol, l.tvol.size()));
    RC(dalloc(&l.d_tstart, l.tstart.size()));
    PSX_HIP(hipMemcpyAsync(l.d_vols, l.vols.data(), sizeof(VolDesc) * l.vols.size(), hipMemcpyHostToDevice, st));
    PSX_HIP(hipMemcpyAsync(l.d_tvol, l.tvol.data(), sizeof(int32_t) * l.tvol.size(), hipMemcpyHostToDevice, st));
    PSX_HIP(hipMemcpyAsync(l.d_tstart, l.tstart.data(), sizeof(int32_t) * l.tstart.size(), hipMemcpyHostToDevice,
                           st));
    return PSX_OK;
}

int env_int(const char *name, int def) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : def;
}

float env_float(const char *name, float def) {
    const char *e = std::getenv(name);
    return e ? (float)std::atof(e) : def;
}

}  // namespace

MGParams mg_params_from_env() {
    MGParams p;
    p.levels = std::max(1, env_int("PSX_MG_LEVELS", p.levels));
    p.nu = std::max(1, env_int("PSX_MG_NU", p.nu));
    p.nu0 = std::max(1, env_int("PSX_MG_NU0", p.nu0));
    p.coarse_max = std::max(3, env_int("PSX_MG_CMAX", p.coarse_max));
    p.rall0 = env_int("PSX_MG_RALL0", p.rall0) != 0;
    p.coarse_deg = std::max(1, env_int("PSX_MG_CDEG", p.coarse_deg));
    p.ratio = std::max(1.5f, env_float("PSX_MG_RATIO", p.ratio));
    p.coarse_ratio = std::max(1.5f, env_float("PSX_MG_CRATIO", p.coarse_ratio));
    p.power_steps = std::max(0, env_int("PSX_MG_POWER", p.power_steps));
    return p;
}

void mg_free(MG &mg) {
    for (auto &l : mg.lv) {
        void *bufs[] = {l.d_vols, l.d_tvol, l.d_tstart, l.A, l.Ah, l.As, l.lam, l.r, l.dA, l.dB};
        for (void *b : bufs)
            if (b) (void)hipFree(b);
        if (l.own_bx) { (void)hipFree(l.b); (void)hipFree(l.x); }
        if (l.own_minv && l.minv) (void)hipFree(l.minv);
    }
    mg = MG();
}

int mg_launch_rall_tiles() { return env_int("PSX_MG_RALL_TILES", 1024); }
int mg_half_stencils() { return env_int("PSX_MG_HALF", 1); }
int mg_launch_uyz() { return env_int("PSX_MG_UYZ", 3); }

int mg_build(hipStream_t st, const Batch &B, int R, const MGParams &prm, MG &mg) {
    static std::atomic<int> generation{0};
    mg_free(mg);
    mg.gen = ++generation;
    mg.prm = prm;
    mg.R = R;
    const int nvol = (int)B.vols.size();
    // number of levels: coarsen while every axis of every volume keeps >= 3
    // coarse nodes and, past two levels, some axis is longer than
    // coarse_max.  Measured
    // (six-case moduli, rtol 1e-5 and 1e-10): a 16^3 coarsest level is best
    // for the 64^3 pack (3 levels; 4 levels are 4-7 % slower) and the 128^3
    // pack (4 levels; 3 are 35-43 % slower, 5 are 4 % slower); the uneven
    // 32^3-96^3 batch keeps 4 (3 are 4 % slower)
    int L = 1;
    {
        std::vector<VolDesc> cur = B.vols;
        while (L < prm.levels) {
            bool can = true;
            int longest = 0;
            for (auto &d : cur) {
                can &= d.nx >= 6 && d.ny >= 6 && d.nz >= 6;
                longest = std::max(longest, std::max(d.nx, std::max(d.ny, d.nz)));
            }
            if (!can || (L >= 2 && longest <= prm.coarse_max)) break;  // two levels where they fit
            for (auto &d : cur) { d.nx = (d.nx + 1) / 2; d.ny = (d.ny + 1) / 2; d.nz = (d.nz + 1) / 2; }
            L++;
        }
    }
    mg.lv.resize(L);
    // level 0: the batch grid
    MGLevel &l0 = mg.lv[0];
    l0.vols = B.vols; l0.tvol = B.tvol; l0.tstart = B.tstart; l0.ntiles = B.ntiles; l0.N = B.Ntot;
    l0.minv = B.d_minv32;
    // coarse geometry
    for (int l = 1; l < L; l++) {
        MGLevel &c = mg.lv[l];
        const MGLevel &f = mg.lv[l - 1];
        int64_t off = 0;
        int tile = 0;
        for (int v = 0; v < nvol; v++) {
            VolDesc d{};
            d.nx = (f.vols[v].nx + 1) / 2; d.ny = (f.vols[v].ny + 1) / 2; d.nz = (f.vols[v].nz + 1) / 2;
            int64_t N = (int64_t)d.nx * d.ny * d.nz;
            d.off = off; d.tile0 = tile; d.ntile = (int)psx::ceil_div(N, TILE);
            for (int k = 0; k < d.ntile; k++) { c.tvol.push_back(v); c.tstart.push_back(k * TILE); }
            tile += d.ntile;
            off += N;
            c.vols.push_back(d);
        }
        c.N = off;
        c.ntiles = tile;
    }
    for (int l = 0; l < L; l++) RC(upload_level(st, mg.lv[l]));
    for (int l = 0; l < L; l++) RC(alloc_level_vectors(mg.lv[l], R, l > 0));
    RC(mg_operators(st, B, mg));
    mg.built = true;
    return PSX_OK;
}

int mg_operators(hipStream_t st, const Batch &B, MG &mg) {
    const MGParams &prm = mg.prm;
    const int L = (int)mg.lv.size();
    const int nvol = (int)B.vols.size();
    MGLevel &l0 = mg.lv[0];
    // operators: A0 assembled in FP64, Galerkin products down the levels
    // Af: the FP64 operator of the level being processed (freed as soon as
    // the next level's Galerkin product exists); freed on every way out
    psx::DeviceBuf<double> Af, pmax_buf;
    double *tmp = nullptr;
    int maxtiles = 0;
    for (auto &l : mg.lv) maxtiles = std::max(maxtiles, l.ntiles);
    RC(dalloc(&tmp, maxtiles));
    pmax_buf.reset(tmp);
    double *const pmax = tmp;
    // Level 0 and the first Galerkin product from the voxel constants and
    // the projected element matrices (no assembled fine stencil: 1.9 kB a
    // node), PSX_MG_GALERKIN_ELEM=0 the assembled stencil and the general
    // product
    const bool elem = env_int("PSX_MG_GALERKIN_ELEM", 1) != 0;
    psx::DeviceBuf<double> d_mab;
    if (elem) {
        const std::vector<double> mab = galerkin_element_tables();
        RC(dalloc(&tmp, mab.size()));
        d_mab.reset(tmp);
        PSX_HIP(hipMemcpyAsync(d_mab.get(), mab.data(), sizeof(double) * mab.size(), hipMemcpyHostToDevice, st));
    } else {
        RC(dalloc(&tmp, (size_t)NSTEN * l0.N));
        Af.reset(tmp);
        hipLaunchKernelGGL(k_assemble_fine, dim3(grid_for(l0.ntiles)), dim3(256), 0, st, l0.d_vols, l0.d_tvol,
                           l0.d_tstart, l0.ntiles, B.d_ka, B.d_kb, Af.get(), l0.N);
    }
    for (int l = 0; l < L; l++) {
        MGLevel &lv = mg.lv[l];
        if (!lv.lam) RC(dalloc(&lv.lam, nvol));
        if (l == 0 && elem) {
            hipLaunchKernelGGL(k_level_prep0_elem, dim3(grid_for(lv.ntiles)), dim3(256), 0, st, lv.d_vols, lv.d_tvol,
                               lv.d_tstart, lv.ntiles, B.d_ka, B.d_kb, (const double *)d_mab.get(), lv.N,
                               (const double *)B.d_minv, pmax);
        } else if (l == 0) {
            hipLaunchKernelGGL(k_level_prep, dim3(grid_for(lv.ntiles)), dim3(256), 0, st, lv.d_vols, lv.d_tvol,
                               lv.d_tstart, lv.ntiles, (const double *)Af.get(), lv.N, (const double *)B.d_minv,
                               (float *)nullptr, (float *)nullptr, pmax);
        } else {
            if (!lv.A) RC(dalloc(&lv.A, (size_t)NSTEN * lv.N));
            if (!lv.minv) {
                RC(dalloc(&lv.minv, (size_t)6 * lv.N));
                lv.own_minv = true;
            }
            hipLaunchKernelGGL(k_level_prep, dim3(grid_for(lv.ntiles)), dim3(256), 0, st, lv.d_vols, lv.d_tvol,
                               lv.d_tstart, lv.ntiles, (const double *)Af.get(), lv.N, (const double *)nullptr, lv.A,
                               lv.minv, pmax);
            if (mg_half_stencils()) {
                if (!lv.Ah) RC(dalloc(&lv.Ah, (size_t)NSTEN * lv.N));
                if (!lv.As) RC(dalloc(&lv.As, (size_t)lv.N));
                hipLaunchKernelGGL(k_half_stencil, dim3((unsigned)psx::ceil_div(lv.N, 256)), dim3(256), 0, st,
                                   (const float *)lv.A, lv.N, lv.Ah, lv.As);
            }
        }
        hipLaunchKernelGGL(k_vol_max, dim3(nvol), dim3(256), 0, st, lv.d_vols, (const double *)pmax, lv.lam);
        PSX_HIP(hipGetLastError());
        if (l + 1 < L) {
            MGLevel &cv = mg.lv[l + 1];
            RC(dalloc(&tmp, (size_t)NSTEN * cv.N));
            psx::DeviceBuf<double> Ac(tmp);
            if (l == 0 && elem) {
                hipLaunchKernelGGL(k_galerkin_elem, dim3(grid_for(cv.ntiles)), dim3(256), 0, st, lv.d_vols,
                                   cv.d_vols, cv.d_tvol, cv.d_tstart, cv.ntiles, B.d_ka, B.d_kb,
                                   (const double *)d_mab.get(), Ac.get(), cv.N);
                PSX_HIP(hipGetLastError());
                PSX_HIP(hipStreamSynchronize(st));
            } else {
                hipLaunchKernelGGL(k_galerkin, dim3(grid_for(cv.ntiles)), dim3(256), 0, st, lv.d_vols, cv.d_vols,
                                   cv.d_tvol, cv.d_tstart, cv.ntiles, (const double *)Af.get(), lv.N, Ac.get(),
                                   cv.N);
            }
            PSX_HIP(hipGetLastError());
            PSX_HIP(hipStreamSynchronize(st));
            Af = std::move(Ac);
        }
    }
    PSX_HIP(hipStreamSynchronize(st));
    Af.reset();
    pmax_buf.reset();
    if (env_int("PSX_MG_DEBUG", 0)) {
        for (int l = 0; l < L; l++) {
            std::vector<float> lam(nvol);
            PSX_HIP(hipMemcpy(lam.data(), mg.lv[l].lam, sizeof(float) * nvol, hipMemcpyDeviceToHost));
            std::fprintf(stderr, "mg level %d: N=%lld vol0 %dx%dx%d lam0=%g\n", l, (long long)mg.lv[l].N,
                         mg.lv[l].vols[0].nx, mg.lv[l].vols[0].ny, mg.lv[l].vols[0].nz, lam[0]);
        }
    }
    // sharpen the bounds: power iteration on D^-1 A per volume and level
    if (prm.power_steps > 0) {
        psx::CallScope scope;
        double *pn = nullptr, *n0 = nullptr, *n1 = nullptr;
        RC(dalloc(&pn, maxtiles));
        scope.buffer(pn);
        RC(dalloc(&n0, nvol));
        scope.buffer(n0);
        RC(dalloc(&n1, nvol));
        scope.buffer(n1);
        for (int l = 0; l < L; l++) {
            MGLevel &lv = mg.lv[l];
            LevelArgs A = args_of(lv, B, prm.ratio);
            const int g = grid_for(lv.ntiles);
            float *va = lv.dA, *vb = lv.dB;
            hipLaunchKernelGGL(k_pow_init, dim3(g), dim3(TILE), 0, st, A, va);
            for (int it = 0; it < prm.power_steps; it++) {
                if (l == 0)
                    hipLaunchKernelGGL(k_pow_step<true>, dim3(g), dim3(TILE), 0, st, A, (const float *)va, vb,
                                       it ? (const double *)n0 : (const double *)nullptr, pn);
                else
                    hipLaunchKernelGGL(k_pow_step<false>, dim3(g), dim3(TILE), 0, st, A, (const float *)va, vb,
                                       it ? (const double *)n0 : (const double *)nullptr, pn);
                hipLaunchKernelGGL(k_vol_norm, dim3(nvol), dim3(256), 0, st, lv.d_vols, (const double *)pn, n1);
                std::swap(va, vb);
                std::swap(n0, n1);
            }
            hipLaunchKernelGGL(k_pow_lam, dim3((unsigned)psx::ceil_div(nvol, 64)), dim3(64), 0, st, nvol,
                               (const double *)n0, lv.lam, prm.power_safety);
            PSX_HIP(hipGetLastError());
        }
        PSX_HIP(hipStreamSynchronize(st));
    }
    if (env_int("PSX_MG_DEBUG", 0)) {
        for (int l = 0; l < L; l++) {
            std::vector<float> lam(nvol);
            PSX_HIP(hipMemcpy(lam.data(), mg.lv[l].lam, sizeof(float) * nvol, hipMemcpyDeviceToHost));
            std::fprintf(stderr, "mg level %d: lam0 after power iteration %g\n", l, lam[0]);
        }
    }
    return PSX_OK;
}

void mg_apply(hipStream_t st, const Batch &B, MG &mg, const SysState *sys, const float *r, float *z, double *prz,
              const CGUpdate *upd) {
    MGLevel &l0 = mg.lv[0];
    const bool only = mg.lv.size() == 1;
    LevelArgs L = args_of(l0, B, only ? mg.prm.coarse_ratio : mg.prm.ratio);
    l0.b = const_cast<float *>(r);
    l0.x = z;
    if (upd)
        hipLaunchKernelGGL(k_cg_update_first, dim3(grid_for(L.ntiles), mg.R), dim3(TILE), 0, st, L, mg.R, sys, upd->X,
                           const_cast<float *>(r), upd->P, upd->Q, upd->prr, x_untouched(mg, 0) ? z : nullptr, l0.r,
                           l0.dA);
    else
        hipLaunchKernelGGL(k_cheb_first, dim3(grid_for(L.ntiles), mg.R), dim3(TILE), 0, st, L, mg.R, sys,
                           (const float *)r, x_untouched(mg, 0) ? z : nullptr, l0.r, l0.dA);
    bool finished = false;
    float *pend = vcycle(st, B, mg, sys, 0, r, prz, &finished);
    if (!finished)
        hipLaunchKernelGGL(k_final0, dim3(grid_for(L.ntiles), mg.R), dim3(TILE), 0, st, L, mg.R, sys, r, z,
                           (const float *)pend, prz);
}

}  // namespace lsmk
