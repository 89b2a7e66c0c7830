// Device code of the D3Q19 two-phase LBM, shared by the exact translation
// unit (lbm.hip: FP64 populations, no FP contraction, the reference's
// operation order) and the fast one (lbm_fast.hip: FP32 storage, FMA
// contraction, reciprocals in place of divisions).  See lbm.hip for the
// data layout.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>

#include "lbm_frac.h"
#include "psx_common.h"

namespace psx_lbm_dev {

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

// The per-direction weight arrays tr, tb and w0 hold one value per lattice
// shell: the rest direction, the six axis links and the twelve diagonals.
// Reading them at wcls(q) keeps three values live instead of nineteen.
__host__ __device__ constexpr int wcls(int q) { return q == 0 ? 0 : (q < 7 ? 1 : 7); }

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

// FP32 pull layout: post-collision populations in AoSoA blocks of 64 nodes,
// block-major: element (direction k, node n) is e = aos(k, n) =
// (n/64)*19*64 + k*64 + n%64, and element e is the float pair at floats 2e
// (T) and 2e + 1 (the colour word s, frac_encode), so one population is one
// 8-byte access.  A node pulls direction q from the node src streams from,
// or from its own opposite slot for a bounce-back link.  With the extended
// colour word the extension bytes follow the 2*19*aos_pad(nl) floats,
// byte e for element e (xbytes(nl) bytes), so a pull reads T, s and x at
// one element index.
constexpr int AOS = 64;
__host__ __device__ __forceinline__ int64_t aos(int k, int64_t n) {
    return (n >> 6) * (Q * AOS) + (int64_t)k * AOS + (n & (AOS - 1));
}
__host__ __device__ __forceinline__ int64_t aosx(int k, int64_t n) { return aos(k, n); }
__host__ __device__ __forceinline__ int64_t aos_pad(int64_t nl) { return (nl + AOS - 1) / AOS * AOS; }
// bytes of extension bytes per population buffer
__host__ __device__ __forceinline__ int64_t xbytes(int64_t nl) { return Q * aos_pad(nl); }
// T and the colour word of element e, and both in one access
template <class I>
__device__ __forceinline__ float2 pop_ld(const float *__restrict__ p, I e) {
    return reinterpret_cast<const float2 *>(p)[e];
}
template <class I>
__device__ __forceinline__ float pop_word(const float *__restrict__ p, I e) {
    return pop_ld(p, e).y;
}
__device__ __forceinline__ void pop_st(float *__restrict__ p, int64_t e, float t, float s) {
    reinterpret_cast<float2 *>(p)[e] = make_float2(t, s);
}

struct PullF32 {
    // the colour field stays FP64: calcFs differentiates it, and an FP32 grid
    // moved the blue permeability by 4e-3 after 20000 steps at 256^3
    using phi_t = double;
    const float *__restrict__ p;
    const int32_t *__restrict__ src;  // [18][nl], -1 = bounce-back
    int64_t nl;
    const uint8_t *__restrict__ x;  // extension bytes of the colour words, or null
    __device__ __forceinline__ void load_c(const LbmConst &c, int q, int64_t i, double &fr, double &fb) const {
        int64_t j = i;
        int k = q;
        if (q > 0) {
            int32_t s = src[(int64_t)(q - 1) * nl + i];
            if (s >= 0) j = s;
            else k = opp(q);
        }
        const float2 ts = pop_ld(p, aos(k, j));
        double f = c.w0[wcls(q)] + (double)ts.x;  // w0 is the same for q and opp(q)
        frac_split(f, ts.y, fr, fb, x ? (uint32_t)x[aosx(k, j)] : 0u);
    }
};

template <class St>
__device__ __forceinline__ void load_pop(const St &s, const LbmConst &c, int q, int64_t i, double &fr,
                                         double &fb) {
    if constexpr (std::is_same<St, PullF32>::value) s.load_c(c, q, i, fr, fb);
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

// sqrt(x) for x >= 0: EXACT the correctly rounded one; else the hardware
// reciprocal square root with one coupled Newton step (relative error near
// 2^-50), exact 0 at 0
template <bool EXACT>
__device__ __forceinline__ double fs_sqrt(double x) {
    if (EXACT) return sqrt(x);
    const double rs = __builtin_amdgcn_rsq(x);
    const double y = x * rs, h = 0.5 * rs;
    const double y1 = fma(y, fma(-y, h, 0.5), y);
    return x > 0.0 ? y1 : 0.0;
}

// calcFs (L1:2626-2796) on the 3x3x3 window of the colour field around a
// voxel; phi3(i, j, k) returns the value at offset (i-1, j-1, k-1).
template <bool EXACT, class Phi3>
__device__ __forceinline__ void calc_fs_window(const LbmConst &c, const Phi3 &phi3, double *Cg, double *Fs) {
    double cc[27];  // cc[i + 3j + 9k]
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double g = phi3(i, j, k);
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
    const double norm2 = Cg[0] * Cg[0] + Cg[1] * Cg[1] + Cg[2] * Cg[2];
    double dn0 = nm[1][0] + nm[3][0] + nm[5][0] + nm[7][0] - nm[0][0] - nm[2][0] - nm[4][0] - nm[6][0];
    double dn1 = nm[3][1] + nm[7][1] + nm[2][1] + nm[6][1] - nm[0][1] - nm[1][1] - nm[4][1] - nm[5][1];
    double dn2 = nm[4][2] + nm[5][2] + nm[6][2] + nm[7][2] - nm[0][2] - nm[1][2] - nm[2][2] - nm[3][2];
    double k1 = dn0 + dn1 + dn2;
    double mm[8];
#pragma unroll
    for (int i = 0; i < 8; i++) mm[i] = fs_sqrt<EXACT>(nm[i][0] * nm[i][0] + nm[i][1] * nm[i][1] + nm[i][2] * nm[i][2]);
    double dm0 = mm[1] + mm[3] + mm[5] + mm[7] - mm[0] - mm[2] - mm[4] - mm[6];
    double dm1 = mm[3] + mm[7] + mm[2] + mm[6] - mm[0] - mm[1] - mm[4] - mm[5];
    double dm2 = mm[4] + mm[5] + mm[6] + mm[7] - mm[0] - mm[1] - mm[2] - mm[3];
    double k2 = Cg[0] * dm0 + Cg[1] * dm1 + Cg[2] * dm2;
    if (norm2 != 0.0) {
        if (EXACT) {
            const double normc = sqrt(norm2);
            k2 = k2 / normc;
            Cg[0] = Cg[0] / normc; Cg[1] = Cg[1] / normc; Cg[2] = Cg[2] / normc;
        } else {  // 1 / |Cg| from the reciprocal square root and two Newton steps
            double in = __builtin_amdgcn_rsq(norm2);
            in = in * fma(-0.5 * norm2 * in, in, 1.5);
            in = in * fma(-0.5 * norm2 * in, in, 1.5);
            k2 *= in;
            Cg[0] *= in; Cg[1] *= in; Cg[2] *= in;
        }
    }
    Fs[0] = c.sig * (k2 - k1) * Cg[0];
    Fs[1] = c.sig * (k2 - k1) * Cg[1];
    Fs[2] = c.sig * (k2 - k1) * Cg[2];
}

// calcFs reading the periodic phi grid (z fastest) around voxel (x,y,z).
template <typename P, bool EXACT = true>
__device__ __forceinline__ void calc_fs(const LbmConst &c, const P *__restrict__ phi, int x, int y, int z,
                                        double *Cg, double *Fs) {
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const int xs[3] = {x == 0 ? nx - 1 : x - 1, x, x == nx - 1 ? 0 : x + 1};
    const int ys[3] = {y == 0 ? ny - 1 : y - 1, y, y == ny - 1 ? 0 : y + 1};
    const int zs[3] = {z == 0 ? nz - 1 : z - 1, z, z == nz - 1 ? 0 : z + 1};
    auto phi3 = [&](int i, int j, int k) -> double {
        return (double)phi[(int64_t)zs[k] + (int64_t)nz * (ys[j] + (int64_t)ny * xs[i])];
    };
    calc_fs_window<EXACT>(c, phi3, Cg, Fs);
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

// Per-node output of save_results (L1:2074-2142): out[0..2] velocity, out[3]
// colour, out[4] pressure, each [nl]; phase = the ind_red_blue code
// (1 red, 2 blue, 6 red side of the interface, 7 blue side).  Both colour
// branches of the reference use the red body force bfr.
template <class St>
__global__ __launch_bounds__(256) void k_fields(const LbmConst c, const St in, const int32_t *__restrict__ vox,
                                                const typename St::phi_t *__restrict__ phi,
                                                double *__restrict__ out, int32_t *__restrict__ phase) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = c.nl;
    if (i >= nl) return;
    double f[Q], rr, rb;
    {
        double a, b;
        load_pop(in, c, 0, i, rr, rb);
        f[0] = rr + rb;
#pragma unroll
        for (int q = 1; q < Q; q++) {
            load_pop(in, c, q, i, a, b);
            f[q] = a + b;
            rr = rr + a;
            rb = rb + b;
        }
    }
    const double cf = (rr - rb) / (rr + rb);
    int x, y, z;
    vox_xyz(c, vox[i], x, y, z);
    double Cg[3], Fs[3], bfi[3];
    calc_fs(c, phi, x, y, z, Cg, Fs);
    const double ri = c.dti / (rr + rb);
#pragma unroll
    for (int d = 0; d < 3; d++) bfi[d] = c.dt * (c.bfr[d] + Fs[d]) * 0.5;
    out[i] = (f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14] + bfi[0]) * ri;
    out[nl + i] = (f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18] + bfi[1]) * ri;
    out[2 * nl + i] = (f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18] + bfi[2]) * ri;
    out[3 * nl + i] = cf;
    out[4 * nl + i] = (rr * c.cr2 + rb * c.cb2) - c.cr2 * 1.0;
    phase[i] = cf > 0.9 ? 1 : cf < -0.9 ? 2 : (cf >= 0.0 && cf <= 0.9) ? 6 : 7;
}

// LSM-LBM coupling force (L1:3849-3879) on the coupled model's solid nodes.
// Every bounce-back link gives -2 (f_out - 0.99 fq) c_in rsi/dt^2 to the 2
// (edge link) or 4 (face link) solid nodes of its wall, half or a quarter
// each.  Two passes: k_link_force evaluates y = -2 (f_out - 0.99 fq) once per
// link (links in node order, so the population reads stay local), then
// k_node_force sums y c rsi/dt^2 x share over the links of every solid node
// that has any, in the reference's loop order (edge links, then face links,
// each in table order): the FP64 path reproduces the reference's sums bit
// for bit.
struct CplLinks {
    const int32_t *lnode;    // [nlink] fluid node (storage index)
    const int8_t *lq;        // [nlink] direction of the returning population (0-based)
    double *y;               // [nlink] per-link term
    const int32_t *start;    // [nsurf+1] contribution range of surface solid node s
    const uint32_t *contrib;  // link id | direction << 27; ids < nedge are edge links
    const int32_t *sgrid;    // [nsurf] node-grid index of surface solid node s
    int32_t nlink, nsurf, nedge, pad_;
    int64_t N;               // node-grid size; af is [3][N]
    double fq[Q];            // rr0 (cr2 dt^2/6 | cr2 dt^2/12), L1:3417
    double cv[Q][3];         // lattice velocities
    double rsdti2;           // rsi/dt^2
};
constexpr int CPL_QSHIFT = 27;

// FOut(n, qb): the post-collision population that returns to node n in
// direction qb, fr + fb.
template <class FOut>
__global__ __launch_bounds__(256) void k_link_force(const CplLinks L, const FOut fout) {
    const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= L.nlink) return;
    const int qb = L.lq[li];
    L.y[li] = -2.0 * (fout(L.lnode[li], qb) - 0.99 * L.fq[qb]);
}

static __global__ __launch_bounds__(256) void k_node_force(const CplLinks L, double *__restrict__ af) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= L.nsurf) return;
    double a[3] = {0.0, 0.0, 0.0};
    for (int32_t k = L.start[s]; k < L.start[s + 1]; k++) {
        const uint32_t e = L.contrib[k];
        const int32_t li = (int32_t)(e & ((1u << CPL_QSHIFT) - 1u));
        const int qb = (int)(e >> CPL_QSHIFT);
        const double y = L.y[li];
        const double share = li < L.nedge ? 0.5 : 0.25;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double floc = y * L.cv[qb][d] * L.rsdti2;
            a[d] = a[d] + floc * share;
        }
    }
    const int64_t g = L.sgrid[s];
#pragma unroll
    for (int d = 0; d < 3; d++) af[d * L.N + g] = a[d];
}

// check_state partial sums; see NDIAG layout in diag_finish().
constexpr int NDIAG = 36;
constexpr int BLOCK = 256;

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

static __global__ __launch_bounds__(256) void k_reduce_cols(const double *__restrict__ partial, int64_t nrows,
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

}  // namespace psx_lbm_dev
