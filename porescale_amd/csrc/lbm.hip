// D3Q19 colour-gradient two-phase LBM for gfx950.
//
// Replaces the time loop body of majorityrule in
// LBM-LSM/lsm_lbm_static_alpha_beta_multipore.f ("L1"): collision
// (L1:2407-2620), calcFs (L1:2626-2796), the bulk/interface branch and
// recolouring (L1:3701-3820), streaming (L1:3833-3839), half-way bounce-back
// (L1:3890-3915) and the density pass (L1:3934-3945); plus check_state
// (L1:2898-3296).
//
// Layout in HBM (nl fluid nodes, reference node order: x outer, z fastest):
//   populations  SoA, direction-major: pop[q][node].  FP64 mode holds fr and
//                fb (2*19 doubles per node).  FP32 mode holds, per (q,node),
//                T = (fr+fb) - tr_q*rr0 and R = fr/(fr+fb) as floats: T keeps
//                the small deviation of the total population (momentum is
//                exact to FP32 ulps of the deviation) and R is exactly 1 or 0
//                for single-colour populations, so pure-phase nodes keep
//                exactly zero minority density as the reference's branch
//                test rb<=ee requires.
//   dst[18][nl]  push target per (direction, node): destination node of the
//                streaming link, or -1 for a bounce-back link (the population
//                returns to the same node in the opposite direction).
//   vox[nl]      voxel index of each node in the phi grid.
//   phi grid     dense nx*ny*nz FP64, z fastest (index z + nz*(y + ny*x)):
//                colour field cf at fluid voxels, `wet` at solid voxels, so
//                the 27-point stencil of calcFs is plain loads.
//
// One time step = two kernels: collide+push (reads own populations, writes
// the streamed post-collision populations into the other buffer) and density
// (reads the new populations, writes phi).  FP64 mode compiles without FP
// contraction and keeps the reference's operation order, so its state is
// bit-identical to L1 after any number of steps.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "psx_common.h"

namespace {

constexpr int Q = 19;
constexpr int H_IV[19][3] = {
    {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {-1, 1, 0},  {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {1, 0, -1},
    {-1, 0, -1}, {0, 1, 1},  {0, -1, 1}, {0, 1, -1},  {0, -1, -1}};
// index of the opposite velocity (niv, L1:515-523)
__host__ __device__ constexpr int opp(int q) {
    return q == 0 ? 0 : (q < 7 ? ((q & 1) ? q + 1 : q - 1) : 7 + 4 * ((q - 7) / 4) + 3 - (q - 7) % 4);
}

struct LbmConst {
    double tr[Q], tb[Q], Sr[Q], Sb[Q], Fpr[Q], Fpb[Q], w0[Q];
    double bfr[3], bfb[3];
    double cr2, cb2, B, sig, wet, ee, dt, dt2, dti;
    int64_t nl;
    int32_t nx, ny, nz, pad_;
};

// ----------------------------------------------------------- storage ----
struct StoreF64 {
    using phi_t = double;
    double *__restrict__ p;
    int64_t nl;
    __device__ __forceinline__ void load(int q, int64_t i, double &fr, double &fb) const {
        fr = p[(int64_t)q * nl + i];
        fb = p[(int64_t)(Q + q) * nl + i];
    }
    __device__ __forceinline__ void store(const LbmConst &, int q, int64_t j, double fr, double fb) const {
        p[(int64_t)q * nl + j] = fr;
        p[(int64_t)(Q + q) * nl + j] = fb;
    }
};

struct StoreF32 {
    using phi_t = double;  // the colour field stays FP64: its differences drive the surface force
    float *__restrict__ p;  // T[19][nl] then R[19][nl]
    int64_t nl;
    __device__ __forceinline__ void load_c(const LbmConst &c, int q, int64_t i, double &fr, double &fb) const {
        double f = c.w0[q] + (double)p[(int64_t)q * nl + i];
        double r = (double)p[(int64_t)(Q + q) * nl + i];
        fr = r * f;
        fb = f - fr;
    }
    __device__ __forceinline__ void store(const LbmConst &c, int q, int64_t j, double fr, double fb) const {
        double f = fr + fb;
        float r;
        if (fb == 0.0) r = 1.0f;
        else if (fr == 0.0) r = 0.0f;
        else if (f == 0.0) r = 0.5f;
        else r = (float)(fr / f);
        p[(int64_t)q * nl + j] = (float)(f - c.w0[q]);
        p[(int64_t)(Q + q) * nl + j] = r;
    }
};

template <class St>
__device__ __forceinline__ void load_pop(const St &s, const LbmConst &c, int q, int64_t i, double &fr,
                                         double &fb) {
    if constexpr (std::is_same<St, StoreF32>::value) s.load_c(c, q, i, fr, fb);
    else s.load(q, i, fr, fb);
}

// ---------------------------------------------------------- collision ----
// collision (L1:2407-2620) on one node; f[0] updated in place, pcf[1..18].
__device__ __forceinline__ void collide_node(double t1, double t2, double t3, double *f, double *pcf,
                                             double r, const double *bf, const double *Fp,
                                             const double *S, double dt, double dti, double dt2) {
    const double c1d6 = 1.0 / 6.0, c1d12 = 1.0 / 12.0, c5d399 = 5.0 / 399.0,
                 c19d399 = 19.0 / 399.0, c11d2394 = 11.0 / 2394.0, c1d63 = 1.0 / 63.0,
                 c1d18 = 1.0 / 18.0, c1d36 = 1.0 / 36.0, c4d1197 = 4.0 / 1197.0,
                 c1d252 = 1.0 / 252.0, c1d72 = 1.0 / 72.0, c1d24 = 1.0 / 24.0;
    double j0 = (f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14]) * dti;
    double j1 = (f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18]) * dti;
    double j2 = (f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18]) * dti;
    double dr1 = 1.0 / r;
    double u0 = (j0 + bf[0] * 0.5 * dt) * dr1;
    double u1 = (j1 + bf[1] * 0.5 * dt) * dr1;
    double u2v = (j2 + bf[2] * 0.5 * dt) * dr1;
    double u1x2 = u0 * u0, u2x2 = u1 * u1, u3x2 = u2v * u2v;
    double u2 = (u1x2 + u2x2 + u3x2) * 0.5;
    double fe[Q];
    fe[0] = r * (t1 - u2 * dt2);
    double j1d6 = j0 * c1d6 * dt, j2d6 = j1 * c1d6 * dt, j3d6 = j2 * c1d6 * dt;
    double rd4 = r * 0.25 * dt2;
    double com1 = (t2 - u2 * c1d6 * dt2) * r;
    double fe1 = com1 + rd4 * u1x2, fe2 = com1 + rd4 * u2x2, fe3 = com1 + rd4 * u3x2;
    fe[1] = fe1 + j1d6; fe[2] = fe1 - j1d6;
    fe[3] = fe2 + j2d6; fe[4] = fe2 - j2d6;
    fe[5] = fe3 + j3d6; fe[6] = fe3 - j3d6;
    double j1p2 = (j0 + j1) * c1d12 * dt, j1m2 = (j0 - j1) * c1d12 * dt;
    double j1p3 = (j0 + j2) * c1d12 * dt, j1m3 = (j0 - j2) * c1d12 * dt;
    double j2p3 = (j1 + j2) * c1d12 * dt, j2m3 = (j1 - j2) * c1d12 * dt;
    double rd8 = r * 0.125 * dt2;
    double a12 = rd8 * (u0 + u1) * (u0 + u1), b12 = rd8 * (u0 - u1) * (u0 - u1);
    double a13 = rd8 * (u0 + u2v) * (u0 + u2v), b13 = rd8 * (u0 - u2v) * (u0 - u2v);
    double a23 = rd8 * (u1 + u2v) * (u1 + u2v), b23 = rd8 * (u1 - u2v) * (u1 - u2v);
    double com2 = (t3 - u2 * c1d12 * dt2) * r;
    fe[7] = com2 + j1p2 + a12;  fe[10] = com2 - j1p2 + a12;
    fe[8] = com2 - j1m2 + b12;  fe[9] = com2 + j1m2 + b12;
    fe[11] = com2 + j1p3 + a13; fe[14] = com2 - j1p3 + a13;
    fe[13] = com2 + j1m3 + b13; fe[12] = com2 - j1m3 + b13;
    fe[15] = com2 + j2p3 + a23; fe[18] = com2 - j2p3 + a23;
    fe[17] = com2 + j2m3 + b23; fe[16] = com2 - j2m3 + b23;
    double fl[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) fl[q] = f[q] - fe[q];
    double com3 = fl[1] + fl[2] + fl[3] + fl[4] + fl[5] + fl[6];
    double com4 = fl[7] + fl[8] + fl[9] + fl[10] + fl[11] + fl[12] + fl[13] + fl[14] + fl[15] +
                  fl[16] + fl[17] + fl[18];
    double mo[Q];
    mo[0] = 0.0;
    mo[1] = -30.0 * fl[0] - 11.0 * com3 + 8.0 * com4;
    mo[2] = 12.0 * fl[0] - 4.0 * com3 + com4;
    mo[3] = 0.0;
    mo[4] = 4.0 * (fl[2] - fl[1]) + fl[7] - fl[8] + fl[9] - fl[10] + fl[11] - fl[12] + fl[13] - fl[14];
    mo[5] = 0.0;
    mo[6] = 4.0 * (fl[4] - fl[3]) + fl[7] + fl[8] - fl[9] - fl[10] + fl[15] - fl[16] + fl[17] - fl[18];
    mo[7] = 0.0;
    mo[8] = 4.0 * (fl[6] - fl[5]) + fl[11] + fl[12] - fl[13] - fl[14] + fl[15] + fl[16] - fl[17] - fl[18];
    double com5 = fl[7] + fl[8] + fl[9] + fl[10];
    double com6 = fl[15] + fl[16] + fl[17] + fl[18];
    double com7 = fl[11] + fl[12] + fl[13] + fl[14];
    double com8 = fl[3] + fl[4] - fl[5] - fl[6];
    double com9 = fl[3] + fl[4] + fl[5] + fl[6];
    mo[9] = 2.0 * (fl[1] + fl[2] - com6) - com9 + com5 + com7;
    mo[10] = -4.0 * (fl[1] + fl[2]) + 2.0 * (com9 - com6) + com5 + com7;
    mo[11] = com8 + com5 - com7;
    mo[12] = -2.0 * com8 + com5 - com7;
    mo[13] = fl[7] - fl[8] - fl[9] + fl[10];
    mo[14] = fl[15] - fl[16] - fl[17] + fl[18];
    mo[15] = fl[11] - fl[12] - fl[13] + fl[14];
    mo[16] = fl[7] - fl[8] + fl[9] - fl[10] - fl[11] + fl[12] - fl[13] + fl[14];
    mo[17] = -fl[7] - fl[8] + fl[9] + fl[10] + fl[15] - fl[16] + fl[17] - fl[18];
    mo[18] = fl[11] + fl[12] - fl[13] - fl[14] - fl[15] - fl[16] + fl[17] + fl[18];
#pragma unroll
    for (int q = 0; q < Q; q++) mo[q] = mo[q] * S[q];
    double mb[Q];
    mb[0] = -c5d399 * mo[1] + c19d399 * mo[2];
    double com10 = -c11d2394 * mo[1] - c1d63 * mo[2];
    mb[1] = com10 - 0.1 * mo[4] + c1d18 * (mo[9] - mo[10]);
    mb[2] = mb[1] + mo[4] * 0.2;
    mb[3] = com10 - 0.1 * mo[6] + c1d36 * (mo[10] - mo[9]) + c1d12 * (mo[11] - mo[12]);
    mb[4] = mb[3] + mo[6] * 0.2;
    mb[5] = com10 - 0.1 * mo[8] + c1d36 * (mo[10] - mo[9]) + c1d12 * (mo[12] - mo[11]);
    mb[6] = mb[5] + mo[8] * 0.2;
    double com11 = c4d1197 * mo[1] + c1d252 * mo[2];
    mb[7] = com11 + (mo[4] + mo[6]) * 0.025 + c1d36 * mo[9] + c1d72 * mo[10] + c1d12 * mo[11] +
            c1d24 * mo[12] + mo[13] * 0.25 + (mo[16] - mo[17]) * 0.125;
    mb[8] = mb[7] - 0.05 * mo[4] - 0.5 * mo[13] - 0.25 * mo[16];
    mb[9] = mb[7] + 0.25 * mo[17] - 0.05 * mo[6] - 0.5 * mo[13];
    mb[10] = mb[9] + 0.5 * mo[13] - 0.05 * mo[4] - 0.25 * mo[16];
    mb[11] = com11 + (mo[4] + mo[8]) * 0.025 + c1d36 * mo[9] + c1d72 * mo[10] - c1d12 * mo[11] -
             c1d24 * mo[12] + mo[15] * 0.25 + (mo[18] - mo[16]) * 0.125;
    mb[12] = mb[11] + 0.25 * mo[16] - 0.5 * mo[15] - 0.05 * mo[4];
    mb[13] = mb[11] - 0.05 * mo[8] - 0.5 * mo[15] - 0.25 * mo[18];
    mb[14] = mb[12] + 0.5 * mo[15] - 0.05 * mo[8] - 0.25 * mo[18];
    mb[15] = com11 + (mo[6] + mo[8]) * 0.025 - c1d18 * mo[9] - c1d36 * mo[10] + mo[14] * 0.25 +
             (mo[17] - mo[18]) * 0.125;
    mb[16] = mb[15] - 0.05 * mo[6] - 0.5 * mo[14] - 0.25 * mo[17];
    mb[17] = mb[15] + 0.25 * mo[18] - 0.05 * mo[8] - 0.5 * mo[14];
    mb[18] = mb[17] + 0.5 * mo[14] - 0.05 * mo[6] - 0.25 * mo[17];
    f[0] = f[0] - mb[0] + Fp[0];
#pragma unroll
    for (int q = 1; q < Q; q++) pcf[q] = f[q] - mb[q] + Fp[q];
}

// calcFs (L1:2626-2796) reading the phi grid around voxel (x,y,z).
template <typename P>
__device__ __forceinline__ void calc_fs(const LbmConst &c, const P *__restrict__ phi, int x, int y, int z,
                                        double *Cg, double *Fs) {
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    int xs[3] = {x == 0 ? nx - 1 : x - 1, x, x == nx - 1 ? 0 : x + 1};
    int ys[3] = {y == 0 ? ny - 1 : y - 1, y, y == ny - 1 ? 0 : y + 1};
    int zs[3] = {z == 0 ? nz - 1 : z - 1, z, z == nz - 1 ? 0 : z + 1};
    double cc[27];  // cc[i + 3j + 9k]
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double g = (double)phi[(int64_t)zs[k] + (int64_t)nz * (ys[j] + (int64_t)ny * xs[i])];
                cc[i + 3 * j + 9 * k] = (g - c.wet) + c.wet;  // solid voxels hold wet
            }
#define C3(i, j, k) cc[((i)-1) + 3 * ((j)-1) + 9 * ((k)-1)]
    double r11 = C3(1,1,1)+C3(2,1,1), r12 = C3(2,1,1)+C3(3,1,1), r13 = C3(1,2,1)+C3(2,2,1),
           r14 = C3(2,2,1)+C3(3,2,1), r15 = C3(1,3,1)+C3(2,3,1), r16 = C3(2,3,1)+C3(3,3,1),
           r17 = C3(1,1,1)+C3(1,2,1), r18 = C3(2,1,1)+C3(2,2,1), r19 = C3(3,1,1)+C3(3,2,1),
           r110 = C3(1,2,1)+C3(1,3,1), r111 = C3(2,2,1)+C3(2,3,1), r112 = C3(3,2,1)+C3(3,3,1);
    double r21 = C3(1,1,2)+C3(2,1,2), r22 = C3(2,1,2)+C3(3,1,2), r23 = C3(1,2,2)+C3(2,2,2),
           r24 = C3(2,2,2)+C3(3,2,2), r25 = C3(1,3,2)+C3(2,3,2), r26 = C3(2,3,2)+C3(3,3,2),
           r27 = C3(1,1,2)+C3(1,2,2), r28 = C3(2,1,2)+C3(2,2,2), r29 = C3(3,1,2)+C3(3,2,2),
           r210 = C3(1,2,2)+C3(1,3,2), r211 = C3(2,2,2)+C3(2,3,2), r212 = C3(3,2,2)+C3(3,3,2);
    double r31 = C3(1,1,3)+C3(2,1,3), r32 = C3(2,1,3)+C3(3,1,3), r33 = C3(1,2,3)+C3(2,2,3),
           r34 = C3(2,2,3)+C3(3,2,3), r35 = C3(1,3,3)+C3(2,3,3), r36 = C3(2,3,3)+C3(3,3,3),
           r37 = C3(1,1,3)+C3(1,2,3), r38 = C3(2,1,3)+C3(2,2,3), r39 = C3(3,1,3)+C3(3,2,3),
           r310 = C3(1,2,3)+C3(1,3,3), r311 = C3(2,2,3)+C3(2,3,3), r312 = C3(3,2,3)+C3(3,3,3);
#undef C3
    double g12 = r18 + r28, g13 = r13 + r23, g34 = r111 + r211, g24 = r14 + r24;
    double g15 = r21 + r23, g37 = r23 + r25, g48 = r24 + r26, g26 = r22 + r24;
    double g56 = r28 + r38, g57 = r23 + r33, g78 = r211 + r311, g68 = r24 + r34;
    double nm[8][3];
    nm[0][0] = g12 - (r17 + r27); nm[1][0] = r19 + r29 - g12;
    nm[2][0] = g34 - (r110 + r210); nm[3][0] = r112 + r212 - g34;
    nm[4][0] = g56 - (r27 + r37); nm[5][0] = r29 + r39 - g56;
    nm[6][0] = g78 - (r210 + r310); nm[7][0] = r212 + r312 - g78;
    nm[0][1] = g13 - (r11 + r21); nm[1][1] = g24 - (r12 + r22);
    nm[2][1] = r15 + r25 - g13; nm[3][1] = r16 + r26 - g24;
    nm[4][1] = g57 - (r21 + r31); nm[5][1] = g68 - (r22 + r32);
    nm[6][1] = r25 + r35 - g57; nm[7][1] = r26 + r36 - g68;
    nm[0][2] = g15 - (r11 + r13); nm[1][2] = g26 - (r12 + r14);
    nm[2][2] = g37 - (r13 + r15); nm[3][2] = g48 - (r14 + r16);
    nm[4][2] = r31 + r33 - g15; nm[5][2] = r32 + r34 - g26;
    nm[6][2] = r33 + r35 - g37; nm[7][2] = r34 + r36 - g48;
#pragma unroll
    for (int d = 0; d < 3; d++)
        Cg[d] = nm[0][d] + nm[1][d] + nm[2][d] + nm[3][d] + nm[4][d] + nm[5][d] + nm[6][d] + nm[7][d];
    double normc = sqrt(Cg[0] * Cg[0] + Cg[1] * Cg[1] + Cg[2] * Cg[2]);
    double dn0 = nm[1][0] + nm[3][0] + nm[5][0] + nm[7][0] - nm[0][0] - nm[2][0] - nm[4][0] - nm[6][0];
    double dn1 = nm[3][1] + nm[7][1] + nm[2][1] + nm[6][1] - nm[0][1] - nm[1][1] - nm[4][1] - nm[5][1];
    double dn2 = nm[4][2] + nm[5][2] + nm[6][2] + nm[7][2] - nm[0][2] - nm[1][2] - nm[2][2] - nm[3][2];
    double k1 = dn0 + dn1 + dn2;
    double mm[8];
#pragma unroll
    for (int i = 0; i < 8; i++) mm[i] = sqrt(nm[i][0] * nm[i][0] + nm[i][1] * nm[i][1] + nm[i][2] * nm[i][2]);
    double dm0 = mm[1] + mm[3] + mm[5] + mm[7] - mm[0] - mm[2] - mm[4] - mm[6];
    double dm1 = mm[3] + mm[7] + mm[2] + mm[6] - mm[0] - mm[1] - mm[4] - mm[5];
    double dm2 = mm[4] + mm[5] + mm[6] + mm[7] - mm[0] - mm[1] - mm[2] - mm[3];
    double k2 = Cg[0] * dm0 + Cg[1] * dm1 + Cg[2] * dm2;
    if (normc != 0.0) {
        k2 = k2 / normc;
        Cg[0] = Cg[0] / normc; Cg[1] = Cg[1] / normc; Cg[2] = Cg[2] / normc;
    }
    Fs[0] = c.sig * (k2 - k1) * Cg[0];
    Fs[1] = c.sig * (k2 - k1) * Cg[1];
    Fs[2] = c.sig * (k2 - k1) * Cg[2];
}

__device__ __forceinline__ void vox_xyz(const LbmConst &c, int32_t v, int &x, int &y, int &z) {
    z = v % c.nz;
    int t = v / c.nz;
    y = t % c.ny;
    x = t / c.ny;
}

// ------------------------------------------------------------ kernels ----
template <class St>
__global__ __launch_bounds__(256) void k_collide(const LbmConst c, const St in, const St out,
                                                 const int32_t *__restrict__ dst,
                                                 const int32_t *__restrict__ vox,
                                                 const typename St::phi_t *__restrict__ phi, int xcd) {
    int64_t b = xcd ? psx::xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    int64_t i = b * blockDim.x + threadIdx.x;
    if (i >= c.nl) return;
    double fr[Q], fb[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) load_pop(in, c, q, i, fr[q], fb[q]);
    double rr = fr[0], rb = fb[0];
#pragma unroll
    for (int q = 1; q < Q; q++) { rr = rr + fr[q]; rb = rb + fb[q]; }
    const double dt = c.dt, dt2 = c.dt2, dti = c.dti;
    double f[Q], pcf[Q], Fp[Q], S[Q], bf[3];
    double t1, t2, t3, r;
    int mode;  // 0 pure red, 1 pure blue, 2 interface (L1:3701-3721)
    double Cg[3] = {0.0, 0.0, 0.0};
    double ti2 = 0.0, ti3 = 0.0;
    if (rb <= c.ee) {
        mode = 0;
        t1 = c.tr[0]; t2 = c.tr[1]; t3 = c.tr[7]; r = rr;
#pragma unroll
        for (int q = 0; q < Q; q++) { f[q] = fr[q]; Fp[q] = c.Fpr[q]; S[q] = c.Sr[q]; }
        bf[0] = c.bfr[0]; bf[1] = c.bfr[1]; bf[2] = c.bfr[2];
    } else if (rr <= c.ee) {
        mode = 1;
        t1 = c.tb[0]; t2 = c.tb[1]; t3 = c.tb[7]; r = rb;
#pragma unroll
        for (int q = 0; q < Q; q++) { f[q] = fb[q]; Fp[q] = c.Fpb[q]; S[q] = c.Sb[q]; }
        bf[0] = c.bfb[0]; bf[1] = c.bfb[1]; bf[2] = c.bfb[2];
    } else {
        mode = 2;
        int x, y, z;
        vox_xyz(c, vox[i], x, y, z);
        double Fs[3];
        calc_fs(c, phi, x, y, z, Cg, Fs);
#pragma unroll
        for (int q = 0; q < Q; q++) f[q] = fr[q] + fb[q];
        r = rr + rb;
        double ci2 = (c.cr2 * rr + c.cb2 * rb) / r;
        t1 = 1.0 - 2.0 * ci2 * dt2;
        t2 = ci2 * (1.0 / 6.0) * dt2;
        t3 = ci2 * (1.0 / 12.0) * dt2;
        ti2 = t2; ti3 = t3;
        bool red = rr >= rb;
#pragma unroll
        for (int d = 0; d < 3; d++) bf[d] = (red ? c.bfr[d] : c.bfb[d]) + Fs[d];
#pragma unroll
        for (int q = 0; q < Q; q++) S[q] = red ? c.Sr[q] : c.Sb[q];
        const double c1d6 = 1.0 / 6.0, c1d12 = 1.0 / 12.0;
        Fp[0] = 0.0;
        Fp[1] = c1d6 * bf[0] * dt2; Fp[2] = -Fp[1];
        Fp[3] = c1d6 * bf[1] * dt2; Fp[4] = -Fp[3];
        Fp[5] = c1d6 * bf[2] * dt2; Fp[6] = -Fp[5];
        Fp[7] = c1d12 * (bf[0] + bf[1]) * dt2;
        Fp[8] = c1d12 * (-bf[0] + bf[1]) * dt2;
        Fp[9] = -Fp[8]; Fp[10] = -Fp[7];
        Fp[11] = c1d12 * (bf[0] + bf[2]) * dt2;
        Fp[12] = c1d12 * (-bf[0] + bf[2]) * dt2;
        Fp[13] = -Fp[12]; Fp[14] = -Fp[11];
        Fp[15] = c1d12 * (bf[1] + bf[2]) * dt2;
        Fp[16] = c1d12 * (-bf[1] + bf[2]) * dt2;
        Fp[17] = -Fp[16]; Fp[18] = -Fp[15];
    }
    collide_node(t1, t2, t3, f, pcf, r, bf, Fp, S, dt, dti, dt2);
    double outr[Q], outb[Q];
    if (mode == 0) {
        outr[0] = f[0];
        outb[0] = c.tb[0] * rb;
#pragma unroll
        for (int q = 1; q < Q; q++) { outr[q] = pcf[q]; outb[q] = c.tb[q] * rb; }
    } else if (mode == 1) {
        outb[0] = f[0];
        outr[0] = c.tr[0] * rr;
#pragma unroll
        for (int q = 1; q < Q; q++) { outb[q] = pcf[q]; outr[q] = c.tr[q] * rr; }
    } else {
        const double cn = 1.0 / sqrt(2.0);
        double cosf[18];
        cosf[0] = Cg[0]; cosf[1] = -cosf[0];
        cosf[2] = Cg[1]; cosf[3] = -cosf[2];
        cosf[4] = Cg[2]; cosf[5] = -cosf[4];
        cosf[6] = (Cg[0] + Cg[1]) * cn; cosf[7] = (-Cg[0] + Cg[1]) * cn;
        cosf[8] = -cosf[7]; cosf[9] = -cosf[6];
        cosf[10] = (Cg[0] + Cg[2]) * cn; cosf[11] = (-Cg[0] + Cg[2]) * cn;
        cosf[12] = -cosf[11]; cosf[13] = -cosf[10];
        cosf[14] = (Cg[1] + Cg[2]) * cn; cosf[15] = (-Cg[1] + Cg[2]) * cn;
        cosf[16] = -cosf[15]; cosf[17] = -cosf[14];
        double ri = r;
        outr[0] = f[0] * rr / ri;
        outb[0] = f[0] * rb / ri;
#pragma unroll
        for (int q = 1; q < Q; q++) {
            double tw = q < 7 ? ti2 : ti3;
            outr[q] = (pcf[q] + c.B * tw * cosf[q - 1] * rb) * rr / ri;
            outb[q] = (pcf[q] - c.B * tw * cosf[q - 1] * rr) * rb / ri;
        }
    }
    out.store(c, 0, i, outr[0], outb[0]);
#pragma unroll
    for (int q = 1; q < Q; q++) {
        int32_t d = dst[(int64_t)(q - 1) * c.nl + i];
        if (d >= 0) out.store(c, q, d, outr[q], outb[q]);
        else out.store(c, opp(q), i, outr[q], outb[q]);
    }
}

template <class St>
__global__ __launch_bounds__(256) void k_density(const LbmConst c, const St in,
                                                 const int32_t *__restrict__ vox,
                                                 typename St::phi_t *__restrict__ phi, int xcd) {
    int64_t b = xcd ? psx::xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    int64_t i = b * blockDim.x + threadIdx.x;
    if (i >= c.nl) return;
    double fr, fb, rr, rb;
    load_pop(in, c, 0, i, rr, rb);
#pragma unroll
    for (int q = 1; q < Q; q++) {
        load_pop(in, c, q, i, fr, fb);
        rr = rr + fr;
        rb = rb + fb;
    }
    phi[vox[i]] = (typename St::phi_t)((rr - rb) / (rr + rb));
}

// check_state partial sums; see NDIAG layout in diag_finish().
constexpr int NDIAG = 36;

template <class St>
__global__ __launch_bounds__(256) void k_diag(const LbmConst c, const St in, const int32_t *__restrict__ vox,
                                              const typename St::phi_t *__restrict__ phi,
                                              double *__restrict__ partial) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[NDIAG];
#pragma unroll
    for (int k = 0; k < NDIAG; k++) v[k] = 0.0;
    if (i < c.nl) {
        double fr[Q], fb[Q];
        int neg = 0;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            load_pop(in, c, q, i, fr[q], fb[q]);
            if (fb[q] < 0.0 || fr[q] < 0.0) neg = 1;
        }
        double rr = fr[0], rb = fb[0];
#pragma unroll
        for (int q = 1; q < Q; q++) { rr = rr + fr[q]; rb = rb + fb[q]; }
        double cf = (rr - rb) / (rr + rb);
        v[0] = rr; v[1] = rb; v[2] = rr + rb;
        if (cf > 0.99) { v[3] = 1.0; v[4] = rr; }
        else if (cf < -0.99) { v[5] = 1.0; v[6] = rb; }
        v[7] = neg; v[8] = rr < 0.0 ? 1.0 : 0.0; v[9] = rb < 0.0 ? 1.0 : 0.0;
        {   // blue momentum as written in L1:3064-3069 (dti binds to the last term)
            const double *g = fb, dti = c.dti;
            v[10] = g[1] - g[2] + g[7] - g[8] + g[9] - g[10] + g[11] - g[12] + g[13] - g[14] * dti;
            v[11] = g[3] - g[4] + g[7] + g[8] - g[9] - g[10] + g[15] - g[16] + g[17] - g[18] * dti;
            v[12] = g[5] - g[6] + g[11] + g[12] - g[13] - g[14] + g[15] + g[16] - g[17] - g[18] * dti;
        }
        int x, y, z;
        vox_xyz(c, vox[i], x, y, z);
        double Cg[3], Fs[3], f[Q], bfi[3], jm[3];
        calc_fs(c, phi, x, y, z, Cg, Fs);
#pragma unroll
        for (int q = 0; q < Q; q++) f[q] = fr[q] + fb[q];
        double ri = 1.0 / (rr + rb);
        bool red = rr >= rb;
#pragma unroll
        for (int d = 0; d < 3; d++) bfi[d] = c.dt * ((red ? c.bfr[d] : c.bfb[d]) + Fs[d]) * 0.5;
        jm[0] = (f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14] + bfi[0]) * c.dti;
        jm[1] = (f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18] + bfi[1]) * c.dti;
        jm[2] = (f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18] + bfi[2]) * c.dti;
        int o = red ? 0 : 3;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            v[13 + o + d] = jm[d];         // Qr / Qb
            v[19 + d] = jm[d] * ri;        // um
            v[22 + o + d] = jm[d] * ri;    // umr / umb
            v[28 + o + d] = bfi[d];        // tfr / tfb
        }
        v[red ? 34 : 35] = 1.0;
    }
    __shared__ double red_s[4][NDIAG];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NDIAG; k++) {
        double s = psx::wave_sum(v[k]);
        if (lane == 0) red_s[w][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NDIAG) {
        int k = threadIdx.x;
        partial[(int64_t)blockIdx.x * NDIAG + k] = ((red_s[0][k] + red_s[1][k]) + red_s[2][k]) + red_s[3][k];
    }
}

__global__ __launch_bounds__(256) void k_reduce_cols(const double *__restrict__ partial, int64_t nrows,
                                                     int ncols, double *__restrict__ out) {
    int k = blockIdx.x;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < nrows; r += blockDim.x) s += partial[r * ncols + k];
    __shared__ double sh[4];
    s = psx::wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[k] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

template <typename P>
__global__ void k_fill(P *p, int64_t n, P v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

// ================================================================ host ====
struct psx_lbm {
    psx_lbm_params p{};
    int nx = 0, ny = 0, nz = 0;
    int64_t nl = 0, nps = 0, npb2 = 0, npb4 = 0, ns = 0;
    std::vector<int32_t> ind_raw, icor, mijk;
    std::vector<int32_t> dst_h;  // [18][nl]
    LbmConst c{};
    hipStream_t stream = nullptr;
    int32_t *d_dst = nullptr, *d_vox = nullptr;
    void *d_pop[2] = {nullptr, nullptr};
    void *d_phi = nullptr;
    double *d_partial = nullptr, *d_red = nullptr;
    int cur = 0;
    int xcd = 1;
    bool f32 = false;
};

namespace {

inline int wrap1(int t, int n) { return t < 1 ? n : (t > n ? 1 : t); }

// Classifies link (node, q) by the rules of L1:1077-1340: returns the
// destination node id (1-based) for a streaming link, 0 for bounce-back.
// Diagonal links stream only if the destination and both single-axis
// projections are fluid (the logical intent of the product test nsum != 0).
int32_t link_dest(const psx_lbm *h, const int32_t *ic, int q) {
    auto M = [&](int x, int y, int z) {
        return h->mijk[(int64_t)(x - 1) + (int64_t)h->nx * ((y - 1) + (int64_t)h->ny * (z - 1))];
    };
    int x = wrap1(ic[0] + H_IV[q][0], h->nx), y = wrap1(ic[1] + H_IV[q][1], h->ny),
        z = wrap1(ic[2] + H_IV[q][2], h->nz);
    int32_t m = M(x, y, z);
    if (m == 0) return 0;
    if (q < 7) return m;
    int32_t m1 = M(wrap1(ic[0] + H_IV[q][0], h->nx), ic[1], ic[2]);
    int32_t m2 = M(ic[0], wrap1(ic[1] + H_IV[q][1], h->ny), ic[2]);
    int32_t m3 = M(ic[0], ic[1], wrap1(ic[2] + H_IV[q][2], h->nz));
    return (m1 != 0 && m2 != 0 && m3 != 0) ? m : 0;
}


// Node ids of the coupled model's solid lattice: corners of ind==0 voxels,
// glued in x, y and z, numbered z outer / x fastest (L1:556-645).  They fill
// columns 3.. of ipb2/ipb4 (L1:1196-1335) and give settings_19's ns.
void coupling_nodes(psx_lbm *h, std::vector<int32_t> &mijks) {
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const int ax = nx + 1, ay = ny + 1, az = nz + 1;
    std::vector<int8_t> inda((size_t)ax * ay * az, 1);
    auto IA = [&](int i, int j, int k) -> int8_t & {
        return inda[(size_t)(i - 1) + (size_t)ax * ((j - 1) + (size_t)ay * (k - 1))];
    };
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                int32_t v = h->ind_raw[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))];
                if ((v > 1 ? 1 : v) == 0)
                    for (int cc = 0; cc < 8; cc++) IA(i + (cc & 1), j + ((cc >> 1) & 1), k + (cc >> 2)) = 0;
            }
    for (int j = 1; j <= ay; j++)
        for (int k = 1; k <= az; k++) IA(1, j, k) = std::min(IA(1, j, k), IA(ax, j, k));
    for (int i = 1; i <= ax; i++)
        for (int k = 1; k <= az; k++) IA(i, 1, k) = std::min(IA(i, 1, k), IA(i, ay, k));
    for (int i = 1; i <= ax; i++)
        for (int j = 1; j <= ay; j++) IA(i, j, 1) = std::min(IA(i, j, 1), IA(i, j, az));
    mijks.assign((size_t)nx * ny * nz, 0);
    int32_t ns = 0;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
                if (IA(i, j, k) == 0) mijks[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))] = ++ns;
    h->ns = ns;
}

void fill_const(psx_lbm *h) {
    const psx_lbm_params &p = h->p;
    LbmConst &c = h->c;
    double dt = p.dt;
    c.cr2 = p.cr2;
    c.cb2 = p.rr0 * p.cr2 / p.rb0;
    double tr1 = 1.0 - 2.0 * p.cr2 * dt * dt, tr2 = p.cr2 * dt * dt * (1.0 / 6.0),
           tr3 = p.cr2 * dt * dt * (1.0 / 12.0);
    double tb1 = 1.0 - 2.0 * c.cb2 * dt * dt, tb2 = c.cb2 * dt * dt * (1.0 / 6.0),
           tb3 = c.cb2 * dt * dt * (1.0 / 12.0);
    for (int q = 0; q < Q; q++) {
        c.tr[q] = q == 0 ? tr1 : (q < 7 ? tr2 : tr3);
        c.tb[q] = q == 0 ? tb1 : (q < 7 ? tb2 : tb3);
        c.w0[q] = c.tr[q] * p.rr0;
    }
    double svr = 1.0 / (3.0 * p.vir * dt + 0.5), svb = 1.0 / (3.0 * p.vib * dt + 0.5);
    auto trt = [&](double sv, double sq, double *S) {
        double v[Q] = {0.0, sv, sv, 0.0, sq, 0.0, sq, 0.0, sq, sv, sv, sv, sv, sv, sv, sv, sq, sq, sq};
        std::memcpy(S, v, sizeof v);
    };
    if (p.collision == PSX_COLL_MRT) {
        std::memcpy(c.Sr, p.Sr, sizeof c.Sr);
        std::memcpy(c.Sb, p.Sb, sizeof c.Sb);
    } else if (p.collision == PSX_COLL_TRT) {
        trt(svr, 1.0 / (p.lambda / (1.0 / svr - 0.5) + 0.5), c.Sr);
        trt(svb, 1.0 / (p.lambda / (1.0 / svb - 0.5) + 0.5), c.Sb);
    } else {  // collmatrTRT (L1:2858-2893)
        trt(svr, 8.0 * (2.0 - svr) / (8.0 - svr), c.Sr);
        trt(svb, 8.0 * (2.0 - svb) / (8.0 - svb), c.Sb);
    }
    double rk[Q];
    for (int q = 0; q < Q; q++) {
        rk[q] = q == 0 ? 1.0 : (q < 7 ? 1.0 / 6.0 : 1.0 / 12.0);
        double sr = 0.0, sb = 0.0;
        for (int d = 0; d < 3; d++) {
            sr += (double)H_IV[q][d] * p.bfr[d];
            sb += (double)H_IV[q][d] * p.bfb[d];
        }
        c.Fpr[q] = rk[q] * sr * (dt * dt);
        c.Fpb[q] = rk[q] * sb * (dt * dt);
    }
    for (int d = 0; d < 3; d++) { c.bfr[d] = p.bfr[d]; c.bfb[d] = p.bfb[d]; }
    c.B = p.B;
    c.sig = 0.03125 * p.sig;
    c.wet = p.wet;
    c.ee = p.ee;
    c.dt = dt;
    c.dt2 = dt * dt;
    c.dti = 1.0 / dt;
    c.nl = h->nl;
    c.nx = h->nx; c.ny = h->ny; c.nz = h->nz;
}

int launch_density(psx_lbm *h, int buf) {
    int64_t nb = psx::ceil_div(h->nl, 256);
    if (nb == 0) return PSX_OK;
    if (h->f32) {
        StoreF32 s{(float *)h->d_pop[buf], h->nl};
        hipLaunchKernelGGL(k_density<StoreF32>, dim3((unsigned)nb), dim3(256), 0, h->stream, h->c, s,
                           h->d_vox, (double *)h->d_phi, h->xcd);
    } else {
        StoreF64 s{(double *)h->d_pop[buf], h->nl};
        hipLaunchKernelGGL(k_density<StoreF64>, dim3((unsigned)nb), dim3(256), 0, h->stream, h->c, s,
                           h->d_vox, (double *)h->d_phi, h->xcd);
    }
    PSX_HIP(hipGetLastError());
    return PSX_OK;
}

int launch_collide(psx_lbm *h) {
    int64_t nb = psx::ceil_div(h->nl, 256);
    if (nb == 0) return PSX_OK;
    int a = h->cur, b = 1 - h->cur;
    if (h->f32) {
        StoreF32 si{(float *)h->d_pop[a], h->nl}, so{(float *)h->d_pop[b], h->nl};
        hipLaunchKernelGGL(k_collide<StoreF32>, dim3((unsigned)nb), dim3(256), 0, h->stream, h->c, si, so,
                           h->d_dst, h->d_vox, (const double *)h->d_phi, h->xcd);
    } else {
        StoreF64 si{(double *)h->d_pop[a], h->nl}, so{(double *)h->d_pop[b], h->nl};
        hipLaunchKernelGGL(k_collide<StoreF64>, dim3((unsigned)nb), dim3(256), 0, h->stream, h->c, si, so,
                           h->d_dst, h->d_vox, (const double *)h->d_phi, h->xcd);
    }
    PSX_HIP(hipGetLastError());
    h->cur = b;
    return PSX_OK;
}

// host-side conversion of one population pair into FP32 storage (same
// formula as StoreF32::store)
inline void f32_encode(double w0, double fr, double fb, float &T, float &R) {
    double f = fr + fb;
    if (fb == 0.0) R = 1.0f;
    else if (fr == 0.0) R = 0.0f;
    else if (f == 0.0) R = 0.5f;
    else R = (float)(fr / f);
    T = (float)(f - w0);
}

}  // namespace

extern "C" int psx_lbm_create(const psx_lbm_params *p, int32_t nx, int32_t ny, int32_t nz,
                              const int32_t *ind, psx_lbm **out) {
    PSX_CHECK_ARG(p && ind && out, "psx_lbm_create: NULL argument");
    PSX_CHECK_ARG(nx >= 3 && ny >= 3 && nz >= 3, "psx_lbm_create: every dimension must be >= 3");
    PSX_CHECK_ARG((int64_t)nx * ny * nz < (int64_t)INT32_MAX, "psx_lbm_create: grid too large");
    PSX_CHECK_ARG(p->dt > 0.0 && p->rr0 > 0.0 && p->rb0 > 0.0, "psx_lbm_create: dt, rr0, rb0 must be > 0");
    PSX_CHECK_ARG(p->idir >= 1 && p->idir <= 3, "psx_lbm_create: idir must be 1, 2 or 3");
    PSX_CHECK_ARG(p->collision >= 0 && p->collision <= 2, "psx_lbm_create: unknown collision model");
    PSX_CHECK_ARG(p->precision == PSX_PREC_F64 || p->precision == PSX_PREC_F32,
                  "psx_lbm_create: unknown precision");
    *out = nullptr;
    int rc = psx::require_device();
    if (rc) return rc;
    psx_lbm *h = nullptr;
    PSX_GUARD_BEGIN
    h = new psx_lbm();
    h->p = *p;
    h->nx = nx; h->ny = ny; h->nz = nz;
    h->f32 = p->precision == PSX_PREC_F32;
    const int64_t nv = (int64_t)nx * ny * nz;
    h->ind_raw.assign(ind, ind + nv);
    h->mijk.assign(nv, 0);
    // fluid list, x outer / z fastest, fluid iff ind == iipore (L1:1046-1072)
    int64_t nl = 0;
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++) {
                int64_t a = (int64_t)(i - 1) + (int64_t)nx * ((j - 1) + (int64_t)ny * (k - 1));
                if (ind[a] == p->iipore) {
                    h->icor.push_back(i); h->icor.push_back(j); h->icor.push_back(k);
                    h->mijk[a] = (int32_t)(++nl);
                }
            }
    h->nl = nl;
    h->dst_h.assign(18 * (size_t)nl, 0);
    std::vector<int32_t> vox(nl);
    for (int64_t n = 0; n < nl; n++) {
        const int32_t *ic = &h->icor[3 * n];
        vox[n] = (int32_t)((ic[2] - 1) + (int64_t)nz * ((ic[1] - 1) + (int64_t)ny * (ic[0] - 1)));
        for (int q = 1; q < Q; q++) {
            int32_t d = link_dest(h, ic, q);
            if (d) { h->nps++; h->dst_h[(size_t)(q - 1) * nl + n] = d - 1; }
            else {
                if (q < 7) h->npb4++; else h->npb2++;
                h->dst_h[(size_t)(q - 1) * nl + n] = -1;
            }
        }
    }
    fill_const(h);
    PSX_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    size_t popb = (size_t)(nl > 0 ? nl : 1) * 2 * Q * (h->f32 ? sizeof(float) : sizeof(double));
    PSX_HIP(hipMalloc(&h->d_pop[0], popb));
    PSX_HIP(hipMalloc(&h->d_pop[1], popb));
    PSX_HIP(hipMemsetAsync(h->d_pop[0], 0, popb, h->stream));
    PSX_HIP(hipMalloc(&h->d_dst, sizeof(int32_t) * 18 * (size_t)(nl > 0 ? nl : 1)));
    PSX_HIP(hipMalloc(&h->d_vox, sizeof(int32_t) * (size_t)(nl > 0 ? nl : 1)));
    size_t phib = (size_t)nv * sizeof(double);
    PSX_HIP(hipMalloc(&h->d_phi, phib));
    int64_t nb = std::max<int64_t>(1, psx::ceil_div(nl, 256));
    PSX_HIP(hipMalloc(&h->d_partial, sizeof(double) * NDIAG * nb));
    PSX_HIP(hipMalloc(&h->d_red, sizeof(double) * NDIAG));
    if (nl > 0) {
        PSX_HIP(hipMemcpyAsync(h->d_dst, h->dst_h.data(), sizeof(int32_t) * 18 * nl, hipMemcpyHostToDevice, h->stream));
        PSX_HIP(hipMemcpyAsync(h->d_vox, vox.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, h->stream));
    }
    int64_t gb = psx::ceil_div(nv, 256);
    hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)gb), dim3(256), 0, h->stream, (double *)h->d_phi, nv, p->wet);
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipStreamSynchronize(h->stream));
    *out = h;
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" void psx_lbm_destroy(psx_lbm *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void *bufs[] = {h->d_pop[0], h->d_pop[1], h->d_dst, h->d_vox, h->d_phi, h->d_partial, h->d_red};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

extern "C" int psx_lbm_counts(psx_lbm *h, int64_t *counts) {
    PSX_CHECK_ARG(h && counts, "psx_lbm_counts: NULL argument");
    // the coupled model's solid node count (settings_19) comes with the tables
    if (h->ns == 0) {
        PSX_GUARD_BEGIN
        std::vector<int32_t> mijks;
        coupling_nodes(h, mijks);
        PSX_GUARD_END
    }
    counts[0] = h->nl; counts[1] = h->nps; counts[2] = h->npb2; counts[3] = h->npb4;
    counts[4] = h->ns;
    return PSX_OK;
}

extern "C" int psx_lbm_tables(psx_lbm *h, int32_t *icor, int32_t *ips, int32_t *ipb2, int32_t *ipb4) {
    PSX_CHECK_ARG(h, "psx_lbm_tables: NULL handle");
    PSX_GUARD_BEGIN
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    if (icor) std::memcpy(icor, h->icor.data(), sizeof(int32_t) * 3 * h->nl);
    if (!ips && !ipb2 && !ipb4) return PSX_OK;
    std::vector<int32_t> mijks;
    coupling_nodes(h, mijks);
    auto MS = [&](int x, int y, int z) {
        return mijks[(size_t)(wrap1(x, nx) - 1) + (size_t)nx * ((wrap1(y, ny) - 1) + (size_t)ny * (wrap1(z, nz) - 1))];
    };
    int64_t s = 0, b2 = 0, b4 = 0;
    for (int64_t n = 0; n < h->nl; n++) {
        const int32_t *ic = &h->icor[3 * n];
        for (int q = 1; q < Q; q++) {
            int32_t d = h->dst_h[(size_t)(q - 1) * h->nl + n];
            const int *v = H_IV[q];
            if (d >= 0) {
                if (ips) { ips[3 * s] = (int32_t)(n + 1); ips[3 * s + 1] = d + 1; ips[3 * s + 2] = q + 1; }
                s++;
            } else if (q < 7) {
                if (ipb4) {
                    int32_t *e = &ipb4[6 * b4];
                    e[0] = (int32_t)(n + 1); e[1] = opp(q) + 1;
                    for (int m = 0; m < 4; m++) {
                        int o1 = m & 1, o2 = (m >> 1) & 1, off[3];
                        if (v[0] != 0) { off[0] = (v[0] + 1) / 2; off[1] = o1; off[2] = o2; }
                        else if (v[1] != 0) { off[0] = o1; off[1] = (v[1] + 1) / 2; off[2] = o2; }
                        else { off[0] = o1; off[1] = o2; off[2] = (v[2] + 1) / 2; }
                        e[2 + m] = MS(ic[0] + off[0], ic[1] + off[1], ic[2] + off[2]);
                    }
                }
                b4++;
            } else {
                if (ipb2) {
                    int32_t *e = &ipb2[4 * b2];
                    e[0] = (int32_t)(n + 1); e[1] = opp(q) + 1;
                    for (int m = 0; m < 2; m++) {
                        int off[3];
                        if (v[0] == 0) { off[0] = m; off[1] = (v[1] + 1) / 2; off[2] = (v[2] + 1) / 2; }
                        else if (v[1] == 0) { off[0] = (v[0] + 1) / 2; off[1] = m; off[2] = (v[2] + 1) / 2; }
                        else { off[0] = (v[0] + 1) / 2; off[1] = (v[1] + 1) / 2; off[2] = m; }
                        e[2 + m] = MS(ic[0] + off[0], ic[1] + off[1], ic[2] + off[2]);
                    }
                }
                b2++;
            }
        }
    }
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_lbm_rates(psx_lbm *h, double *Sr, double *Sb) {
    PSX_CHECK_ARG(h && Sr && Sb, "psx_lbm_rates: NULL argument");
    std::memcpy(Sr, h->c.Sr, sizeof h->c.Sr);
    std::memcpy(Sb, h->c.Sb, sizeof h->c.Sb);
    return PSX_OK;
}

extern "C" int psx_lbm_set_state(psx_lbm *h, const double *fr, const double *fb) {
    PSX_CHECK_ARG(h && fr && fb, "psx_lbm_set_state: NULL argument");
    if (h->nl == 0) return PSX_OK;
    PSX_GUARD_BEGIN
    const int64_t nl = h->nl;
    h->cur = 0;
    if (h->f32) {
        std::vector<float> st((size_t)2 * Q * nl);
        for (int64_t i = 0; i < nl; i++)
            for (int q = 0; q < Q; q++)
                f32_encode(h->c.w0[q], fr[Q * i + q], fb[Q * i + q], st[(size_t)q * nl + i],
                           st[(size_t)(Q + q) * nl + i]);
        PSX_HIP(hipMemcpyAsync(h->d_pop[0], st.data(), sizeof(float) * st.size(), hipMemcpyHostToDevice, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
    } else {
        std::vector<double> st((size_t)2 * Q * nl);
        for (int64_t i = 0; i < nl; i++)
            for (int q = 0; q < Q; q++) {
                st[(size_t)q * nl + i] = fr[Q * i + q];
                st[(size_t)(Q + q) * nl + i] = fb[Q * i + q];
            }
        PSX_HIP(hipMemcpyAsync(h->d_pop[0], st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
    }
    int rc = launch_density(h, 0);
    if (rc) return rc;
    PSX_HIP(hipStreamSynchronize(h->stream));
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_lbm_get_state(psx_lbm *h, double *fr, double *fb) {
    PSX_CHECK_ARG(h && fr && fb, "psx_lbm_get_state: NULL argument");
    if (h->nl == 0) return PSX_OK;
    PSX_GUARD_BEGIN
    const int64_t nl = h->nl;
    if (h->f32) {
        std::vector<float> st((size_t)2 * Q * nl);
        PSX_HIP(hipMemcpyAsync(st.data(), h->d_pop[h->cur], sizeof(float) * st.size(), hipMemcpyDeviceToHost, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
        for (int64_t i = 0; i < nl; i++)
            for (int q = 0; q < Q; q++) {
                double f = h->c.w0[q] + (double)st[(size_t)q * nl + i];
                double r = (double)st[(size_t)(Q + q) * nl + i];
                fr[Q * i + q] = r * f;
                fb[Q * i + q] = f - r * f;
            }
    } else {
        std::vector<double> st((size_t)2 * Q * nl);
        PSX_HIP(hipMemcpyAsync(st.data(), h->d_pop[h->cur], sizeof(double) * st.size(), hipMemcpyDeviceToHost, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
        for (int64_t i = 0; i < nl; i++)
            for (int q = 0; q < Q; q++) {
                fr[Q * i + q] = st[(size_t)q * nl + i];
                fb[Q * i + q] = st[(size_t)(Q + q) * nl + i];
            }
    }
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_lbm_get_fields(psx_lbm *h, double *rr, double *rb, double *cf) {
    PSX_CHECK_ARG(h, "psx_lbm_get_fields: NULL handle");
    PSX_GUARD_BEGIN
    std::vector<double> fr((size_t)Q * h->nl), fb((size_t)Q * h->nl);
    int rc = psx_lbm_get_state(h, fr.data(), fb.data());
    if (rc) return rc;
    for (int64_t i = 0; i < h->nl; i++) {
        double a = fr[Q * i], b = fb[Q * i];
        for (int q = 1; q < Q; q++) { a = a + fr[Q * i + q]; b = b + fb[Q * i + q]; }
        if (rr) rr[i] = a;
        if (rb) rb[i] = b;
        if (cf) cf[i] = (a - b) / (a + b);
    }
    return PSX_OK;
    PSX_GUARD_END
}

extern "C" int psx_lbm_init_random(psx_lbm *h, double sat) {
    PSX_CHECK_ARG(h, "psx_lbm_init_random: NULL handle");
    PSX_GUARD_BEGIN
    // initdistr icon=0 (L1:2822-2841): Park-Miller draws in node order
    std::vector<double> fr((size_t)Q * h->nl), fb((size_t)Q * h->nl);
    double seed = 13.081984;
    for (int64_t i = 0; i < h->nl; i++) {
        seed = std::fmod(16807.0 * seed, 2147483647.0);
        double rn = seed / 2147483648.0;
        double cfield = rn <= sat ? 0.0 : 1.0;
        for (int q = 0; q < Q; q++) {
            fr[Q * i + q] = h->c.tr[q] * cfield * h->p.rr0;
            fb[Q * i + q] = h->c.tb[q] * (1.0 - cfield) * h->p.rb0;
        }
    }
    return psx_lbm_set_state(h, fr.data(), fb.data());
    PSX_GUARD_END
}

extern "C" int psx_lbm_step(psx_lbm *h, int64_t nsteps) {
    PSX_CHECK_ARG(h && nsteps >= 0, "psx_lbm_step: bad argument");
    for (int64_t s = 0; s < nsteps; s++) {
        int rc = launch_collide(h);
        if (rc) return rc;
        rc = launch_density(h, h->cur);
        if (rc) return rc;
    }
    return PSX_OK;
}

extern "C" int psx_lbm_sync(psx_lbm *h) {
    PSX_CHECK_ARG(h, "psx_lbm_sync: NULL handle");
    PSX_HIP(hipStreamSynchronize(h->stream));
    return PSX_OK;
}

extern "C" int psx_lbm_profile(psx_lbm *h, int64_t nsteps, double *ms) {
    PSX_CHECK_ARG(h && ms && nsteps > 0, "psx_lbm_profile: bad argument");
    std::vector<hipEvent_t> ev(2 * nsteps + 1);
    for (auto &e : ev) PSX_HIP(hipEventCreate(&e));
    PSX_HIP(hipEventRecord(ev[0], h->stream));
    for (int64_t s = 0; s < nsteps; s++) {
        int rc = launch_collide(h);
        if (rc) return rc;
        PSX_HIP(hipEventRecord(ev[2 * s + 1], h->stream));
        rc = launch_density(h, h->cur);
        if (rc) return rc;
        PSX_HIP(hipEventRecord(ev[2 * s + 2], h->stream));
    }
    PSX_HIP(hipEventSynchronize(ev[2 * nsteps]));
    double tc = 0.0, td = 0.0;
    for (int64_t s = 0; s < nsteps; s++) {
        float a = 0.f, b = 0.f;
        PSX_HIP(hipEventElapsedTime(&a, ev[2 * s], ev[2 * s + 1]));
        PSX_HIP(hipEventElapsedTime(&b, ev[2 * s + 1], ev[2 * s + 2]));
        tc += a; td += b;
    }
    float tot = 0.f;
    PSX_HIP(hipEventElapsedTime(&tot, ev[0], ev[2 * nsteps]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    ms[0] = tot;
    ms[1] = tc / (double)nsteps;
    ms[2] = td / (double)nsteps;
    return PSX_OK;
}

extern "C" int psx_lbm_footprint(psx_lbm *h, int64_t *bytes) {
    PSX_CHECK_ARG(h && bytes, "psx_lbm_footprint: NULL argument");
    *bytes = (int64_t)2 * Q * (h->f32 ? 4 : 8);
    return PSX_OK;
}

extern "C" int psx_lbm_check_state(psx_lbm *h, int64_t it1, int64_t itw, psx_lbm_diag *d) {
    PSX_CHECK_ARG(h && d, "psx_lbm_check_state: NULL argument");
    std::memset(d, 0, sizeof *d);
    double s[NDIAG] = {0};
    int64_t nb = psx::ceil_div(h->nl, 256);
    if (nb > 0) {
        if (h->f32) {
            StoreF32 st{(float *)h->d_pop[h->cur], h->nl};
            hipLaunchKernelGGL(k_diag<StoreF32>, dim3((unsigned)nb), dim3(256), 0, h->stream, h->c, st, h->d_vox,
                               (const double *)h->d_phi, h->d_partial);
        } else {
            StoreF64 st{(double *)h->d_pop[h->cur], h->nl};
            hipLaunchKernelGGL(k_diag<StoreF64>, dim3((unsigned)nb), dim3(256), 0, h->stream, h->c, st, h->d_vox,
                               (const double *)h->d_phi, h->d_partial);
        }
        PSX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_reduce_cols, dim3(NDIAG), dim3(256), 0, h->stream, h->d_partial, nb, NDIAG, h->d_red);
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipMemcpyAsync(s, h->d_red, sizeof s, hipMemcpyDeviceToHost, h->stream));
        PSX_HIP(hipStreamSynchronize(h->stream));
    }
    const psx_lbm_params &p = h->p;
    const double dti = 1.0 / p.dt;
    double svr = 1.0 / (3.0 * p.vir * p.dt + 0.5), svb = 1.0 / (3.0 * p.vib * p.dt + 0.5);
    double vir = dti * (1.0 / svr - 0.5) / 3.0, vib = dti * (1.0 / svb - 0.5) / 3.0;  // L1:3419-3420
    d->tmr = s[0]; d->tmb = s[1]; d->tm = s[2];
    d->Cpr = p.cr2 * s[4] / s[3];
    d->Cpb = h->c.cb2 * s[6] / s[5];
    d->isneg = s[7] > 0; d->negdens_r = s[8] > 0; d->negdens_b = s[9] > 0;
    double t0 = (double)((int32_t)(h->nx * h->ny * h->nz));
    for (int k = 0; k < 3; k++) {
        d->Qr[k] = s[13 + k] / t0; d->Qb[k] = s[16 + k] / t0; d->um[k] = s[19 + k] / t0;
        d->umr[k] = s[22 + k]; d->umb[k] = s[25 + k]; d->tfr[k] = s[28 + k]; d->tfb[k] = s[31 + k];
    }
    d->nr = s[34]; d->nb = s[35];
    int id = p.idir - 1;
    d->permr = d->Qr[id] * vir / p.bfr[id];
    d->permb = d->Qb[id] * vib / p.bfb[id];
    d->Re = p.lsize * s[12] / (s[1] * vir);
    double nl = (double)h->nl, sig = h->c.sig;
    d->outm[0] = (double)(it1 * itw);
    d->outm[1] = d->permr;
    d->outm[2] = d->permb;
    d->outm[3] = 0.03125 * d->Qr[id] * vir / sig;
    d->outm[4] = 0.03125 * d->Qb[id] * vib / sig;
    d->outm[5] = 0.03125 * d->umr[id] * vir * p.rr0 / (d->nr * sig);
    d->outm[6] = 0.03125 * d->umb[id] * vib * p.rb0 / (d->nb * sig);
    d->outm[7] = d->Cpr - d->Cpb;
    d->outm[8] = d->Cpr;
    d->outm[9] = d->Cpb;
    d->outm[10] = d->nr / nl;
    d->outm[11] = d->nb / nl;
    return PSX_OK;
}
