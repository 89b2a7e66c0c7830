// FP32-storage variant of the two-phase LBM step.  Populations are stored
// per (direction, node) as
//   T = (fr + fb) - tr_q*rr0   (FP32)   and   s = the colour word of the
//   minority colour's fraction of fr + fb (25-bit significand, lbm_frac.h)
// in 32 bits each, side by side as one 8-byte element (arithmetic in FP64
// registers), post-collision and not yet streamed, in AoSoA blocks of 64
// nodes (lbm_device.h: aos()).  Streaming is a pull: each node gathers
// direction q from its source node (src table) or, for a bounce-back link,
// from its own opposite slot, and writes its own slots contiguously.
// Measured on MI355X, gathered reads with contiguous writes move the step's
// populations 1.6x faster than the scattered writes of a push, and the
// 64-node blocks keep each wavefront's 19 element streams inside one 9.7 KB
// span.  This translation unit allows FMA contraction and
// replaces the per-node divisions of the recolouring step (L1:3798-3816) and
// of calcFs by one reciprocal each.
#include <cstdlib>
#include <type_traits>

#include "lbm_fast.h"

using namespace psx_lbm_dev;

namespace {

// x component of c_q, as a compile-time function of q
__host__ __device__ constexpr int cxq(int q) {
    return q == 1 || q == 7 || q == 9 || q == 11 || q == 13 ? 1 : q == 2 || q == 8 || q == 10 || q == 12 || q == 14 ? -1 : 0;
}

// Pull table in one of four forms (template parameter M):
//   M = 2: src = i + d, d the int16 offset (fits when an x-plane holds fewer
//          than ~32k fluid nodes);
//   M = 1: src = i + pb[cx_q + 1] + d, d the int16 rank difference of source
//          and node within their x-planes and pb the per-plane bases of the
//          node's plane;
//   in both d = -32768 marks a bounce-back link;
//   M = 0: the int32 source index (-1: bounce-back), direction-major [18][nl];
//   M = 3: the same node-major, srca[20 i + q - 1] (two pad entries): five
//          16-byte loads per node instead of 18, which leaves a wavefront
//          room for more population loads in flight.
struct PullIdx {
    const int32_t *__restrict__ src;
    const int16_t *__restrict__ src16;
    const int32_t *__restrict__ pbase;  // [3][nx]
    const int32_t *__restrict__ srca = nullptr;  // [nl][20] (M = 3)
    // element table [nl][20] of the collide kernel (collide_pull_e), or null
    const uint32_t *__restrict__ etab = nullptr;
};

// the 256-node block that workgroup blockIdx.x works on
__device__ __forceinline__ int64_t work_block(int xcd) {
    return xcd ? psx::xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
}

template <int M>
__device__ __forceinline__ void load_bases(const LbmConst &c, const PullIdx &P, int32_t v, int32_t (&pb)[3]) {
    if (M == 1) {
        const int X = v / (c.ny * c.nz);
        pb[0] = P.pbase[X]; pb[1] = P.pbase[c.nx + X]; pb[2] = P.pbase[2 * c.nx + X];
    }
}

// slot (direction k, node j) that direction q of node i pulls from
template <int M>
__device__ __forceinline__ void pull_from(const PullIdx &P, const int32_t (&pb)[3], int64_t nl, int64_t i, int q,
                                          int64_t &j, int &k) {
    j = i;
    k = q;
    if (q == 0) return;
    if (M == 1) {
        int d = P.src16[(int64_t)(q - 1) * nl + i];
        if (d == -32768) k = opp(q);
        else j = i + pb[cxq(q) + 1] + d;
    } else if (M == 2) {
        int d = P.src16[(int64_t)(q - 1) * nl + i];
        if (d == -32768) k = opp(q);
        else j = i + d;
    } else {
        int32_t s = P.src[(int64_t)(q - 1) * nl + i];
        if (s >= 0) j = s;
        else k = opp(q);
    }
}

// 1 / x for a normal, finite x: the hardware reciprocal and two Newton
// steps (error below 2^-100 before the final rounding)
__device__ __forceinline__ double frcp(double x) {
    double rc = __builtin_amdgcn_rcp(x);
    rc = fma(rc, fma(-x, rc, 1.0), rc);
    return fma(rc, fma(-x, rc, 1.0), rc);
}

// The colour word of fraction r (frac_word_t) with the common case first:
// for 2^-126 <= r < 1 (after rounding) the word and extension byte are the
// double's own rounded bits repacked, six integer operations; subnormal
// fractions and escapes (parts of opposite sign) take the full conversion,
// which a wavefront then runs only when one of its lanes needs it.
template <int XB>
__device__ __forceinline__ uint32_t frac_word_fast(double r, uint32_t &x) {
    const uint64_t rb = frac_dbits(r) + (1ull << (27 - XB));
    const uint32_t hi = (uint32_t)(rb >> 32), lo = (uint32_t)rb;
    if (hi - (897u << 20) < (126u << 20)) {
        x = XB ? (lo >> (28 - XB)) & ((1u << XB) - 1u) : 0u;
        return (hi << 5) - (896u << 25) + ((lo >> 27) & 0x1Eu);
    }
    return frac_word_t<XB>(r, x);
}

// Stores the population (fr, fb) of direction q of node j: T and the colour
// word, and for X the word's extension byte (33-bit fraction).  The
// minority's fraction m / f is the hardware reciprocal, one Newton step and
// one residual correction: within an ulp of the correctly rounded quotient
// (which costs twice the instructions), so the word's rounding to 25 or 33
// bits is unbiased as before; f near zero or denormal takes the division.
template <bool X>
__device__ __forceinline__ void put(const LbmConst &c, float *__restrict__ out, uint8_t *__restrict__ outx, int q,
                                    int64_t j, double fr, double fb) {
    const double f = fr + fb;
    const bool blue = fabs(fb) < fabs(fr);
    const double mn = blue ? fb : fr;
    double r;
    if (fabs(f) >= 0x1p-1000) {
        double rc = __builtin_amdgcn_rcp(f);
        rc = fma(rc, fma(-f, rc, 1.0), rc);
        const double q0 = mn * rc;
        r = fma(fma(-q0, f, mn), rc, q0);
    } else {
        r = f == 0.0 ? 0.5 : mn / f;
    }
    uint32_t x;
    const uint32_t w = frac_word_fast<X ? 8 : 0>(r, x) | (blue ? 1u : 0u);
    pop_st(out, aos(q, j), (float)(f - c.w0[wcls(q)]), __uint_as_float(w));
    if (X) outx[aosx(q, j)] = (uint8_t)x;
}

// The collision of L1:2407-2620 with the relaxation vector selected per lane
// (red: Sr, else Sb) and the force projection Fp_q = rk_q (c_q.bf) dt^2
// formed in place, updated in place: f[0] becomes the post-collision rest
// population and f[1..18] the post-collision populations.  dr1 = 1 / r.
__device__ __forceinline__ void collide_fast(const LbmConst &c, double t1, double t2, double t3, double (&f)[Q],
                                             double r, double dr1, const double (&bf)[3], bool red) {
    const double dt = c.dt, dti = c.dti, dt2 = c.dt2;
    const double c1d6 = 1.0 / 6.0, c1d12 = 1.0 / 12.0, c5d399 = 5.0 / 399.0,
                 c19d399 = 19.0 / 399.0, c11d2394 = 11.0 / 2394.0, c1d63 = 1.0 / 63.0,
                 c1d18 = 1.0 / 18.0, c1d36 = 1.0 / 36.0, c4d1197 = 4.0 / 1197.0,
                 c1d252 = 1.0 / 252.0, c1d72 = 1.0 / 72.0, c1d24 = 1.0 / 24.0;
    double j0 = (f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14]) * dti;
    double j1 = (f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18]) * dti;
    double j2 = (f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18]) * dti;
    double u0 = (j0 + bf[0] * 0.5 * dt) * dr1;
    double u1 = (j1 + bf[1] * 0.5 * dt) * dr1;
    double u2v = (j2 + bf[2] * 0.5 * dt) * dr1;
    double u1x2 = u0 * u0, u2x2 = u1 * u1, u3x2 = u2v * u2v;
    double u2 = (u1x2 + u2x2 + u3x2) * 0.5;
    double fl[Q];
    fl[0] = f[0] - r * (t1 - u2 * dt2);
    double j1d6 = j0 * c1d6 * dt, j2d6 = j1 * c1d6 * dt, j3d6 = j2 * c1d6 * dt;
    double rd4 = r * 0.25 * dt2;
    double com1 = (t2 - u2 * c1d6 * dt2) * r;
    double fe1 = com1 + rd4 * u1x2, fe2 = com1 + rd4 * u2x2, fe3 = com1 + rd4 * u3x2;
    fl[1] = f[1] - (fe1 + j1d6); fl[2] = f[2] - (fe1 - j1d6);
    fl[3] = f[3] - (fe2 + j2d6); fl[4] = f[4] - (fe2 - j2d6);
    fl[5] = f[5] - (fe3 + j3d6); fl[6] = f[6] - (fe3 - j3d6);
    double j1p2 = (j0 + j1) * c1d12 * dt, j1m2 = (j0 - j1) * c1d12 * dt;
    double j1p3 = (j0 + j2) * c1d12 * dt, j1m3 = (j0 - j2) * c1d12 * dt;
    double j2p3 = (j1 + j2) * c1d12 * dt, j2m3 = (j1 - j2) * c1d12 * dt;
    double rd8 = r * 0.125 * dt2;
    double a12 = rd8 * (u0 + u1) * (u0 + u1), b12 = rd8 * (u0 - u1) * (u0 - u1);
    double a13 = rd8 * (u0 + u2v) * (u0 + u2v), b13 = rd8 * (u0 - u2v) * (u0 - u2v);
    double a23 = rd8 * (u1 + u2v) * (u1 + u2v), b23 = rd8 * (u1 - u2v) * (u1 - u2v);
    double com2 = (t3 - u2 * c1d12 * dt2) * r;
    fl[7] = f[7] - (com2 + j1p2 + a12);  fl[10] = f[10] - (com2 - j1p2 + a12);
    fl[8] = f[8] - (com2 - j1m2 + b12);  fl[9] = f[9] - (com2 + j1m2 + b12);
    fl[11] = f[11] - (com2 + j1p3 + a13); fl[14] = f[14] - (com2 - j1p3 + a13);
    fl[13] = f[13] - (com2 + j1m3 + b13); fl[12] = f[12] - (com2 - j1m3 + b13);
    fl[15] = f[15] - (com2 + j2p3 + a23); fl[18] = f[18] - (com2 - j2p3 + a23);
    fl[17] = f[17] - (com2 + j2m3 + b23); fl[16] = f[16] - (com2 - j2m3 + b23);
    double com3 = fl[1] + fl[2] + fl[3] + fl[4] + fl[5] + fl[6];
    double com4 = fl[7] + fl[8] + fl[9] + fl[10] + fl[11] + fl[12] + fl[13] + fl[14] + fl[15] +
                  fl[16] + fl[17] + fl[18];
    double com5 = fl[7] + fl[8] + fl[9] + fl[10];
    double com6 = fl[15] + fl[16] + fl[17] + fl[18];
    double com7 = fl[11] + fl[12] + fl[13] + fl[14];
    double com8 = fl[3] + fl[4] - fl[5] - fl[6];
    double com9 = fl[3] + fl[4] + fl[5] + fl[6];
#define SEL(q) (red ? c.Sr[q] : c.Sb[q])
    double m1 = (-30.0 * fl[0] - 11.0 * com3 + 8.0 * com4) * SEL(1);
    double m2 = (12.0 * fl[0] - 4.0 * com3 + com4) * SEL(2);
    double m4 = (4.0 * (fl[2] - fl[1]) + fl[7] - fl[8] + fl[9] - fl[10] + fl[11] - fl[12] + fl[13] - fl[14]) * SEL(4);
    double m6 = (4.0 * (fl[4] - fl[3]) + fl[7] + fl[8] - fl[9] - fl[10] + fl[15] - fl[16] + fl[17] - fl[18]) * SEL(6);
    double m8 = (4.0 * (fl[6] - fl[5]) + fl[11] + fl[12] - fl[13] - fl[14] + fl[15] + fl[16] - fl[17] - fl[18]) * SEL(8);
    double m9 = (2.0 * (fl[1] + fl[2] - com6) - com9 + com5 + com7) * SEL(9);
    double m10 = (-4.0 * (fl[1] + fl[2]) + 2.0 * (com9 - com6) + com5 + com7) * SEL(10);
    double m11 = (com8 + com5 - com7) * SEL(11);
    double m12 = (-2.0 * com8 + com5 - com7) * SEL(12);
    double m13 = (fl[7] - fl[8] - fl[9] + fl[10]) * SEL(13);
    double m14 = (fl[15] - fl[16] - fl[17] + fl[18]) * SEL(14);
    double m15 = (fl[11] - fl[12] - fl[13] + fl[14]) * SEL(15);
    double m16 = (fl[7] - fl[8] + fl[9] - fl[10] - fl[11] + fl[12] - fl[13] + fl[14]) * SEL(16);
    double m17 = (-fl[7] - fl[8] + fl[9] + fl[10] + fl[15] - fl[16] + fl[17] - fl[18]) * SEL(17);
    double m18 = (fl[11] + fl[12] - fl[13] - fl[14] - fl[15] - fl[16] + fl[17] + fl[18]) * SEL(18);
#undef SEL
    // mo(1), mo(4), mo(6), mo(8) are zero (conserved moments)
    const double F6 = c1d6 * dt2, F12 = c1d12 * dt2;
    f[0] = f[0] - (-c5d399 * m1 + c19d399 * m2);
    double com10 = -c11d2394 * m1 - c1d63 * m2;
    double mb1 = com10 - 0.1 * m4 + c1d18 * (m9 - m10);
    double mb3 = com10 - 0.1 * m6 + c1d36 * (m10 - m9) + c1d12 * (m11 - m12);
    double mb5 = com10 - 0.1 * m8 + c1d36 * (m10 - m9) + c1d12 * (m12 - m11);
    f[1] = f[1] - mb1 + F6 * bf[0];
    f[2] = f[2] - (mb1 + m4 * 0.2) - F6 * bf[0];
    f[3] = f[3] - mb3 + F6 * bf[1];
    f[4] = f[4] - (mb3 + m6 * 0.2) - F6 * bf[1];
    f[5] = f[5] - mb5 + F6 * bf[2];
    f[6] = f[6] - (mb5 + m8 * 0.2) - F6 * bf[2];
    double com11 = c4d1197 * m1 + c1d252 * m2;
    double mb7 = com11 + (m4 + m6) * 0.025 + c1d36 * m9 + c1d72 * m10 + c1d12 * m11 + c1d24 * m12 + m13 * 0.25 +
                 (m16 - m17) * 0.125;
    double mb8 = mb7 - 0.05 * m4 - 0.5 * m13 - 0.25 * m16;
    double mb9 = mb7 + 0.25 * m17 - 0.05 * m6 - 0.5 * m13;
    double mb10 = mb9 + 0.5 * m13 - 0.05 * m4 - 0.25 * m16;
    double mb11 = com11 + (m4 + m8) * 0.025 + c1d36 * m9 + c1d72 * m10 - c1d12 * m11 - c1d24 * m12 + m15 * 0.25 +
                  (m18 - m16) * 0.125;
    double mb12 = mb11 + 0.25 * m16 - 0.5 * m15 - 0.05 * m4;
    double mb13 = mb11 - 0.05 * m8 - 0.5 * m15 - 0.25 * m18;
    double mb14 = mb12 + 0.5 * m15 - 0.05 * m8 - 0.25 * m18;
    double mb15 = com11 + (m6 + m8) * 0.025 - c1d18 * m9 - c1d36 * m10 + m14 * 0.25 + (m17 - m18) * 0.125;
    double mb16 = mb15 - 0.05 * m6 - 0.5 * m14 - 0.25 * m17;
    double mb17 = mb15 + 0.25 * m18 - 0.05 * m8 - 0.5 * m14;
    double mb18 = mb17 + 0.5 * m14 - 0.05 * m6 - 0.25 * m17;
    f[7] = f[7] - mb7 + F12 * (bf[0] + bf[1]);
    f[10] = f[10] - mb10 - F12 * (bf[0] + bf[1]);
    f[8] = f[8] - mb8 + F12 * (-bf[0] + bf[1]);
    f[9] = f[9] - mb9 - F12 * (-bf[0] + bf[1]);
    f[11] = f[11] - mb11 + F12 * (bf[0] + bf[2]);
    f[14] = f[14] - mb14 - F12 * (bf[0] + bf[2]);
    f[12] = f[12] - mb12 + F12 * (-bf[0] + bf[2]);
    f[13] = f[13] - mb13 - F12 * (-bf[0] + bf[2]);
    f[15] = f[15] - mb15 + F12 * (bf[1] + bf[2]);
    f[18] = f[18] - mb18 - F12 * (bf[1] + bf[2]);
    f[16] = f[16] - mb16 + F12 * (-bf[1] + bf[2]);
    f[17] = f[17] - mb17 - F12 * (-bf[1] + bf[2]);
}

// The node's collision, recolouring and store after its pull: f[q] holds the
// pulled totals and rr, rb the colour densities.  resplit(blue) replaces
// every f[q] by its red (false) or blue (true) part, for the single-colour
// branches; fs(Cg, Fs) evaluates calcFs at the node's voxel.
template <bool X, class Resplit, class CalcFs>
__device__ __forceinline__ void collide_node_f32(const LbmConst &c, double (&f)[Q], double rr, double rb,
                                                 float *__restrict__ out, uint8_t *__restrict__ outx, int64_t i,
                                                 const Resplit &resplit, const CalcFs &fs) {
    const double dt2 = c.dt2;
    double bf[3];
    double t1, t2, t3, r, rinv;
    int mode;  // 0 pure red, 1 pure blue, 2 interface (L1:3701-3721)
    bool red;
    double Cg[3] = {0.0, 0.0, 0.0};
    if (rb <= c.ee) {
        mode = 0; red = true;
        t1 = c.tr[0]; t2 = c.tr[1]; t3 = c.tr[7]; r = rr;
        resplit(false);
        bf[0] = c.bfr[0]; bf[1] = c.bfr[1]; bf[2] = c.bfr[2];
        rinv = frcp(r);
    } else if (rr <= c.ee) {
        mode = 1; red = false;
        t1 = c.tb[0]; t2 = c.tb[1]; t3 = c.tb[7]; r = rb;
        resplit(true);
        bf[0] = c.bfb[0]; bf[1] = c.bfb[1]; bf[2] = c.bfb[2];
        rinv = frcp(r);
    } else {
        mode = 2;
        double Fs[3];
        fs(Cg, Fs);
        r = rr + rb;
        rinv = frcp(r);
        double ci2 = (c.cr2 * rr + c.cb2 * rb) * rinv;
        t1 = 1.0 - 2.0 * ci2 * dt2;
        t2 = ci2 * (1.0 / 6.0) * dt2;
        t3 = ci2 * (1.0 / 12.0) * dt2;
        red = rr >= rb;
#pragma unroll
        for (int d = 0; d < 3; d++) bf[d] = (red ? c.bfr[d] : c.bfb[d]) + Fs[d];
    }
    collide_fast(c, t1, t2, t3, f, r, rinv, bf, red);
    double ar = 0.0, ab = 0.0;
    if (mode == 2) {
        ar = rr * rinv;
        ab = rb * rinv;
    }
    double o_r, o_b;
    if (mode == 0) { o_r = f[0]; o_b = c.tb[0] * rb; }
    else if (mode == 1) { o_b = f[0]; o_r = c.tr[0] * rr; }
    else { o_r = f[0] * ar; o_b = f[0] * ab; }
    put<X>(c, out, outx, 0, i, o_r, o_b);
    const double cn = 0.70710678118654752440;
    const double t2B = t2 * c.B, t3B = t3 * c.B;
#pragma unroll
    for (int q = 1; q < Q; q++) {
        if (mode == 0) { o_r = f[q]; o_b = c.tb[wcls(q)] * rb; }
        else if (mode == 1) { o_b = f[q]; o_r = c.tr[wcls(q)] * rr; }
        else {
            // cos of the angle between the colour gradient and c_q (L1:3786-3803)
            const int cx = H_IV[q][0], cy = H_IV[q][1], cz = H_IV[q][2];
            double cs = cx * Cg[0] + cy * Cg[1] + cz * Cg[2];
            if (q >= 7) cs *= cn;
            double tw = (q < 7 ? t2B : t3B) * cs;
            o_r = (f[q] + tw * rb) * ar;
            o_b = (f[q] - tw * rr) * ab;
        }
        put<X>(c, out, outx, q, i, o_r, o_b);
    }
}

// Pull + collide + recolour of node i (voxel v) into its own slots of out.
// fs(x, y, z, Cg, Fs) evaluates calcFs at the node's voxel.  X: extended
// colour words, their bytes in inx / outx.
template <int M, class CalcFs, bool WIDE = false, bool X = false>
__device__ __forceinline__ void collide_pull(const LbmConst &c, const float *__restrict__ in,
                                             const uint8_t *__restrict__ inx, float *__restrict__ out,
                                             uint8_t *__restrict__ outx, const PullIdx &P, int64_t i, int32_t v,
                                             const CalcFs &fs) {
    const int64_t nl = c.nl;
    int32_t pb[3];
    load_bases<M>(c, P, v, pb);
    int32_t sa[20];  // M = 3: the node's source indices
    if (M == 3) {
        const int4 *p4 = reinterpret_cast<const int4 *>(P.srca + 20 * i);
#pragma unroll
        for (int g = 0; g < 5; g++) {
            const int4 w = p4[g];
            sa[4 * g] = w.x; sa[4 * g + 1] = w.y; sa[4 * g + 2] = w.z; sa[4 * g + 3] = w.w;
        }
    }
    double f[Q];
    // where direction q is pulled from: an AoSoA element index, 32-bit while
    // the population buffer holds fewer than 2^31 floats (nl < 5.6e7), 64-bit
    // beyond (WIDE: 19 more registers)
    typename std::conditional<WIDE, int64_t, int32_t>::type at[Q];
    uint32_t xs[Q];  // extension bytes (X)
    double rr = 0.0, rb = 0.0;
    bool allred = true, allblue = true, esc = false;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        int64_t j;
        int k;
        if (M == 3) {
            j = i;
            k = q;
            if (q > 0) {
                if (sa[q - 1] >= 0) j = sa[q - 1];
                else k = opp(q);
            }
        } else {
            pull_from<M>(P, pb, nl, i, q, j, k);
        }
        at[q] = (decltype(at[0] + 0))aos(k, j);
        const float2 ts = pop_ld(in, at[q]);
        double t = c.w0[wcls(q)] + (double)ts.x;
        const float s = ts.y;
        xs[q] = X ? (uint32_t)inx[at[q]] : 0u;
        double a, b;
        frac_split_fast(t, s, xs[q], a, b, esc);
        f[q] = t;
        rr += a;
        rb += b;
        allred &= frac_all_red(s) && xs[q] == 0u;
        allblue &= frac_all_blue(s) && xs[q] == 0u;
    }
    if (esc) {  // a fraction outside [0, 1) (parts of opposite sign): rare
        rr = 0.0;
        rb = 0.0;
        for (int q = 0; q < Q; q++) {
            double a, b;
            frac_split(f[q], pop_word(in, at[q]), a, b, xs[q]);
            rr += a;
            rb += b;
        }
    }
    auto resplit = [&](bool blue) {  // the colour split of a node that is not single-coloured
        if (blue ? allblue : allred) return;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            double a, b;
            frac_split(f[q], pop_word(in, at[q]), a, b, xs[q]);
            f[q] = blue ? b : a;
        }
    };
    auto fsv = [&](double *Cg, double *Fs) {
        int x, y, z;
        vox_xyz(c, v, x, y, z);
        fs(x, y, z, Cg, Fs);
    };
    collide_node_f32<X>(c, f, rr, rb, out, outx, i, resplit, fsv);
}

// calcFs at voxel v = z + nz (y + ny x) of the periodic FP64 colour grid,
// the neighbours' element offsets formed from (x, y, z) without divisions.
__device__ __forceinline__ void calc_fs_xyz(const LbmConst &c, const double *__restrict__ phi, uint32_t v, int x,
                                           int y, int z, double *Cg, double *Fs) {
    const int nz = c.nz, nyz = c.ny * c.nz;
    const int ox[3] = {x == 0 ? (c.nx - 1) * nyz : -nyz, 0, x == c.nx - 1 ? -(c.nx - 1) * nyz : nyz};
    const int oy[3] = {y == 0 ? (c.ny - 1) * nz : -nz, 0, y == c.ny - 1 ? -(c.ny - 1) * nz : nz};
    const int oz[3] = {z == 0 ? nz - 1 : -1, 0, z == nz - 1 ? -(nz - 1) : 1};
    auto phi3 = [&](int a, int b2, int e) -> double { return phi[v + (uint32_t)(ox[a] + oy[b2] + oz[e])]; };
    calc_fs_window<false>(c, phi3, Cg, Fs);
}

// Pull + collide with the element table: etab[20 i + q - 1] (q = 1..18) is
// the AoSoA element index node i pulls direction q from (the opposite slot
// of its own node for a bounce-back link), etab[20 i + 18] its voxel and
// etab[20 i + 19] its coordinates x | y << 10 | z << 20.  Five 16-byte loads
// give every address; the rare re-splits load the table row again rather
// than hold it in registers.
template <bool X>
__device__ __forceinline__ void collide_pull_e(const LbmConst &c, const float *__restrict__ in,
                                               const uint8_t *__restrict__ inx, float *__restrict__ out,
                                               uint8_t *__restrict__ outx, const uint32_t *__restrict__ etab,
                                               const double *__restrict__ phi, int64_t i) {
    uint32_t e[20];
    const uint4 *p4 = reinterpret_cast<const uint4 *>(etab + 20 * i);
#pragma unroll
    for (int g = 0; g < 5; g++) {
        const uint4 w = p4[g];
        e[4 * g] = w.x; e[4 * g + 1] = w.y; e[4 * g + 2] = w.z; e[4 * g + 3] = w.w;
    }
    const uint32_t self = (uint32_t)aos(0, i);
    // element at as a 32-bit byte offset (the table is built only for
    // buffers below 4 GiB): the load takes the buffer base from scalar
    // registers
    auto ld = [&](uint32_t at) -> float2 {
        return *reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(in) + (at << 3));
    };
    double f[Q];
    double rr = 0.0, rb = 0.0;
    bool esc = false;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const uint32_t at = q == 0 ? self : e[q - 1];
        const float2 ts = ld(at);
        double t = c.w0[wcls(q)] + (double)ts.x;
        const float s = ts.y;
        const uint32_t xb = X ? (uint32_t)inx[at] : 0u;
        double a, b;
        frac_split_fast(t, s, xb, a, b, esc);
        f[q] = t;
        rr += a;
        rb += b;
    }
    // the table row again, for the rare branches below
    auto at_of = [&](int q) -> uint32_t {
        return q == 0 ? self : __builtin_nontemporal_load(etab + 20 * i + q - 1);
    };
    if (esc) {  // a fraction outside [0, 1) (parts of opposite sign): rare
        rr = 0.0;
        rb = 0.0;
        for (int q = 0; q < Q; q++) {
            const uint32_t at = at_of(q);
            double a, b;
            frac_split(f[q], ld(at).y, a, b, X ? (uint32_t)inx[at] : 0u);
            rr += a;
            rb += b;
        }
    }
    auto resplit = [&](bool blue) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const uint32_t at = at_of(q);
            double a, b;
            frac_split(f[q], ld(at).y, a, b, X ? (uint32_t)inx[at] : 0u);
            f[q] = blue ? b : a;
        }
    };
    const uint32_t v = e[18], pxyz = e[19];
    auto fsv = [&](double *Cg, double *Fs) {
        calc_fs_xyz(c, phi, v, (int)(pxyz & 1023u), (int)((pxyz >> 10) & 1023u), (int)(pxyz >> 20), Cg, Fs);
    };
    collide_node_f32<X>(c, f, rr, rb, out, outx, i, resplit, fsv);
}

// Colour field of node i from its pulled populations.
template <int M, bool X = false>
__device__ __forceinline__ double colour_pull(const LbmConst &c, const float *__restrict__ in,
                                              const uint8_t *__restrict__ inx, const PullIdx &P, int64_t i,
                                              int32_t v) {
    int32_t pb[3];
    load_bases<M>(c, P, v, pb);
    double rr = 0.0, rb = 0.0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        int64_t j;
        int k;
        pull_from<M>(P, pb, c.nl, i, q, j, k);
        const int64_t a = aos(k, j);
        const float2 ts = pop_ld(in, a);
        double t = c.w0[wcls(q)] + (double)ts.x;
        double fr, fb;
        frac_split(t, ts.y, fr, fb, X ? (uint32_t)inx[a] : 0u);
        rr += fr;
        rb += fb;
    }
    return (rr - rb) / (rr + rb);
}

// Colour field of node i from the node-major int16 table (direct offsets,
// or plane-based with PB): the same pulls and sums as colour_pull.
template <bool PB, bool X>
__global__ __launch_bounds__(BLOCK) void k_density_f32n(const LbmConst c, const float *__restrict__ in,
                                                        const uint8_t *__restrict__ inx,
                                                        const int32_t *__restrict__ pbase,
                                                        const uint16_t *__restrict__ dtab, double *__restrict__ phi,
                                                        int xcd) {
    const int64_t i = work_block(xcd) * blockDim.x + threadIdx.x;
    if (i >= c.nl) return;
    const uint4 *pa = reinterpret_cast<const uint4 *>(dtab) + 2 * i;
    const uint4 a0 = pa[0], a1 = pa[1];
    const uint2 b = reinterpret_cast<const uint2 *>(dtab + 16 * c.nl)[i];
    const uint32_t w[9] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w, b.x};
    const int32_t v = (int32_t)b.y;
    int32_t pb[3] = {0, 0, 0};
    if (PB) {
        const int Xp = v / (c.ny * c.nz);
        pb[0] = pbase[Xp]; pb[1] = pbase[c.nx + Xp]; pb[2] = pbase[2 * c.nx + Xp];
    }
    double rr = 0.0, rb = 0.0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        int64_t j = i;
        int k = q;
        if (q > 0) {
            const uint32_t h = w[(q - 1) >> 1];
            const int d = (int16_t)((q - 1) & 1 ? h >> 16 : h & 0xFFFFu);
            if (d == -32768) k = opp(q);
            else j = i + (PB ? pb[cxq(q) + 1] : 0) + d;
        }
        const int64_t a = aos(k, j);
        const float2 ts = pop_ld(in, a);
        double t = c.w0[wcls(q)] + (double)ts.x;
        double fr, fb;
        frac_split(t, ts.y, fr, fb, X ? (uint32_t)inx[a] : 0u);
        rr += fr;
        rb += fb;
    }
    phi[v] = (rr - rb) / (rr + rb);
}

template <int M, bool WIDE, bool X>
__global__ __launch_bounds__(BLOCK) void k_collide_f32(const LbmConst c, const float *__restrict__ in,
                                                       const uint8_t *__restrict__ inx, float *__restrict__ out,
                                                       uint8_t *__restrict__ outx, const PullIdx P,
                                                       const int32_t *__restrict__ vox,
                                                       const double *__restrict__ phi, int xcd) {
    int64_t i = work_block(xcd) * blockDim.x + threadIdx.x;
    if (i >= c.nl) return;
    auto fs = [&](int x, int y, int z, double *Cg, double *Fs) { calc_fs<double, false>(c, phi, x, y, z, Cg, Fs); };
    collide_pull<M, decltype(fs), WIDE, X>(c, in, inx, out, outx, P, i, vox[i], fs);
}

template <bool X>
__global__ __launch_bounds__(BLOCK) void k_collide_f32e(const LbmConst c, const float *__restrict__ in,
                                                        const uint8_t *__restrict__ inx, float *__restrict__ out,
                                                        uint8_t *__restrict__ outx, const PullIdx P,
                                                        const double *__restrict__ phi, int xcd) {
    int64_t i = work_block(xcd) * blockDim.x + threadIdx.x;
    if (i >= c.nl) return;
    collide_pull_e<X>(c, in, inx, out, outx, P.etab, phi, i);
}

template <int M, bool X>
__global__ __launch_bounds__(BLOCK) void k_density_f32(const LbmConst c, const float *__restrict__ in,
                                                       const uint8_t *__restrict__ inx, const PullIdx P,
                                                       const int32_t *__restrict__ vox, double *__restrict__ phi,
                                                       int xcd) {
    int64_t i = work_block(xcd) * blockDim.x + threadIdx.x;
    if (i >= c.nl) return;
    const int32_t v = vox[i];
    phi[v] = colour_pull<M, X>(c, in, inx, P, i, v);
}

// One whole step over one brick of voxels per workgroup: the colour field of
// every fluid voxel of the brick and its one-voxel halo goes into LDS (an
// (bx+2) x (by+2) x (bz+2) box, z fastest, solid cells holding wet), then
// the brick's own nodes collide reading the colour window from LDS.  Nodes
// are numbered brick by brick, so a brick's nodes are one index range.
__global__ __launch_bounds__(BLOCK) void k_step_f32(const LbmConst c, const float *__restrict__ in,
                                                    float *__restrict__ out, const int32_t *__restrict__ src,
                                                    const int32_t *__restrict__ vox,
                                                    const psx_lbm_fast::Brick *__restrict__ bricks,
                                                    const int32_t *__restrict__ hnode,
                                                    const int16_t *__restrict__ hcell, int lx, int ly, int lz,
                                                    int xcd) {
    extern __shared__ double phs[];
    const int64_t t = xcd ? psx::xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const psx_lbm_fast::Brick B = bricks[t];
    const PullIdx P{src, nullptr, nullptr};
    const int ncell = lx * ly * lz;
    for (int k = threadIdx.x; k < ncell; k += BLOCK) phs[k] = c.wet;
    __syncthreads();
    for (int k = threadIdx.x; k < B.hcount; k += BLOCK) {
        const int64_t i = hnode[B.hstart + k];
        phs[hcell[B.hstart + k]] = colour_pull<0>(c, in, nullptr, P, i, vox[i]);
    }
    __syncthreads();
    auto fs = [&](int x, int y, int z, double *Cg, double *Fs) {
        const int cell = (z - B.z0) + lz * ((y - B.y0) + ly * (x - B.x0));  // the cell of offset (-1,-1,-1)
        auto phi3 = [&](int a, int b2, int e) -> double { return phs[cell + e + lz * (b2 + ly * a)]; };
        calc_fs_window<false>(c, phi3, Cg, Fs);
    };
    for (int k = threadIdx.x; k < B.ncount; k += BLOCK) {
        const int64_t i = B.nstart + k;
        collide_pull<0>(c, in, nullptr, out, nullptr, P, i, vox[i], fs);
    }
}

}  // namespace

namespace psx_lbm_fast {

bool wide_index(int64_t nl) {
    const char *e = std::getenv("PSX_LBM_WIDE");  // 1: 64-bit pull offsets at any size (testing)
    return (e && std::atoi(e) != 0) || aos_pad(nl > 0 ? nl : 1) * 2 * Q >= (int64_t)INT32_MAX;
}

namespace {
template <int M, bool X>
void launch_collide(const LbmConst &c, const float *in, const uint8_t *inx, float *out, uint8_t *outx,
                    const PullIdx &P, const int32_t *vox, const double *phi, int xcd, hipStream_t st, int64_t nb) {
    if (wide_index(c.nl))
        hipLaunchKernelGGL((k_collide_f32<M, true, X>), dim3((unsigned)nb), dim3(BLOCK), 0, st, c, in, inx, out, outx,
                           P, vox, phi, xcd);
    else
        hipLaunchKernelGGL((k_collide_f32<M, false, X>), dim3((unsigned)nb), dim3(BLOCK), 0, st, c, in, inx, out,
                           outx, P, vox, phi, xcd);
}
template <int M>
void launch_density(const LbmConst &c, const float *in, const uint8_t *inx, const PullIdx &P, const int32_t *vox,
                    double *phi, int xcd, hipStream_t st, int64_t nb) {
    if (inx)
        hipLaunchKernelGGL((k_density_f32<M, true>), dim3((unsigned)nb), dim3(BLOCK), 0, st, c, in, inx, P, vox, phi,
                           xcd);
    else
        hipLaunchKernelGGL((k_density_f32<M, false>), dim3((unsigned)nb), dim3(BLOCK), 0, st, c, in, inx, P, vox,
                           phi, xcd);
}
}  // namespace

hipError_t collide(const LbmConst &c, const float *in, const uint8_t *inx, float *out, uint8_t *outx,
                   const int32_t *src, const int16_t *src16, const int32_t *pbase, const int32_t *srca,
                   const uint32_t *etab, const int32_t *vox, const double *phi, int xcd, hipStream_t st) {
    int64_t nb = psx::ceil_div(c.nl, BLOCK);
    if (nb == 0) return hipSuccess;
    if (etab) {  // the element table: every address precomputed (see collide_pull_e)
        PullIdx P{src, nullptr, nullptr, nullptr};
        P.etab = etab;
        if (inx)
            hipLaunchKernelGGL((k_collide_f32e<true>), dim3((unsigned)nb), dim3(BLOCK), 0, st, c, in, inx, out, outx,
                               P, phi, xcd);
        else
            hipLaunchKernelGGL((k_collide_f32e<false>), dim3((unsigned)nb), dim3(BLOCK), 0, st, c, in, inx, out,
                               outx, P, phi, xcd);
        return hipGetLastError();
    }
    // the collide kernel is latency-bound (occupancy 3): it takes the direct
    // int16 offsets when they exist, else the node-major int32 table (five
    // loads a node, no dependent load of a plane base), else the
    // direction-major one
    if (src16 && !pbase) {
        const PullIdx P{src, src16, nullptr};
        if (inx) launch_collide<2, true>(c, in, inx, out, outx, P, vox, phi, xcd, st, nb);
        else launch_collide<2, false>(c, in, inx, out, outx, P, vox, phi, xcd, st, nb);
    } else if (srca) {
        const PullIdx P{src, nullptr, nullptr, srca};
        if (inx) launch_collide<3, true>(c, in, inx, out, outx, P, vox, phi, xcd, st, nb);
        else launch_collide<3, false>(c, in, inx, out, outx, P, vox, phi, xcd, st, nb);
    } else {
        const PullIdx P{src, nullptr, nullptr, nullptr};
        if (inx) launch_collide<0, true>(c, in, inx, out, outx, P, vox, phi, xcd, st, nb);
        else launch_collide<0, false>(c, in, inx, out, outx, P, vox, phi, xcd, st, nb);
    }
    return hipGetLastError();
}

hipError_t density(const LbmConst &c, const float *in, const uint8_t *inx, const int32_t *src,
                   const int16_t *src16, const int32_t *pbase, const uint16_t *dtab, const int32_t *vox,
                   double *phi, int xcd, hipStream_t st) {
    int64_t nb = psx::ceil_div(c.nl, BLOCK);
    if (nb == 0) return hipSuccess;
    if (dtab) {
        if (pbase && inx)
            hipLaunchKernelGGL((k_density_f32n<true, true>), dim3((unsigned)nb), dim3(BLOCK), 0, st, c, in, inx,
                               pbase, dtab, phi, xcd);
        else if (pbase)
            hipLaunchKernelGGL((k_density_f32n<true, false>), dim3((unsigned)nb), dim3(BLOCK), 0, st, c, in, inx,
                               pbase, dtab, phi, xcd);
        else if (inx)
            hipLaunchKernelGGL((k_density_f32n<false, true>), dim3((unsigned)nb), dim3(BLOCK), 0, st, c, in, inx,
                               pbase, dtab, phi, xcd);
        else
            hipLaunchKernelGGL((k_density_f32n<false, false>), dim3((unsigned)nb), dim3(BLOCK), 0, st, c, in, inx,
                               pbase, dtab, phi, xcd);
        return hipGetLastError();
    }
    const PullIdx P{src, src16, pbase};
    if (src16 && pbase) launch_density<1>(c, in, inx, P, vox, phi, xcd, st, nb);
    else if (src16) launch_density<2>(c, in, inx, P, vox, phi, xcd, st, nb);
    else launch_density<0>(c, in, inx, P, vox, phi, xcd, st, nb);
    return hipGetLastError();
}

hipError_t step(const LbmConst &c, const float *in, float *out, const int32_t *src, const int32_t *vox,
                const Brick *bricks, int nbricks, const int32_t *hnode, const int16_t *hcell, const int (&box)[3],
                int xcd, hipStream_t st) {
    if (nbricks == 0) return hipSuccess;
    const size_t lds = sizeof(double) * (size_t)box[0] * box[1] * box[2];
    hipLaunchKernelGGL(k_step_f32, dim3((unsigned)nbricks), dim3(BLOCK), lds, st, c, in, out, src, vox, bricks,
                       hnode, hcell, box[0], box[1], box[2], xcd);
    return hipGetLastError();
}

namespace {
// post-collision population returning in direction qb: stored, unstreamed,
// in slot (opp(qb), n)
struct FOutF32 {
    const float *p;
    double w0[Q];
    __device__ __forceinline__ double operator()(int64_t n, int qb) const {
        const int q = opp(qb);
        return w0[q] + (double)p[2 * aos(q, n)];
    }
};
}  // namespace

hipError_t link_force(const LbmConst &c, const float *pop, const CplLinks &L, hipStream_t st) {
    if (L.nlink == 0) return hipSuccess;
    FOutF32 f{pop, {}};
    for (int q = 0; q < Q; q++) f.w0[q] = c.w0[q];
    hipLaunchKernelGGL(k_link_force<FOutF32>, dim3((unsigned)psx::ceil_div(L.nlink, 256)), dim3(256), 0, st, L, f);
    return hipGetLastError();
}

hipError_t diag(const LbmConst &c, const float *in, const uint8_t *inx, const int32_t *src, const int32_t *vox,
                const double *phi, double *partial, hipStream_t st) {
    int64_t nb = psx::ceil_div(c.nl, BLOCK);
    if (nb == 0) return hipSuccess;
    PullF32 ld{in, src, c.nl, inx};
    hipLaunchKernelGGL(k_diag<PullF32>, dim3((unsigned)nb), dim3(BLOCK), 0, st, c, ld, vox, phi, partial);
    return hipGetLastError();
}

hipError_t fields(const LbmConst &c, const float *in, const uint8_t *inx, const int32_t *src, const int32_t *vox,
                  const double *phi, double *out, int32_t *phase, hipStream_t st) {
    int64_t nb = psx::ceil_div(c.nl, BLOCK);
    if (nb == 0) return hipSuccess;
    PullF32 ld{in, src, c.nl, inx};
    hipLaunchKernelGGL(k_fields<PullF32>, dim3((unsigned)nb), dim3(BLOCK), 0, st, c, ld, vox, phi, out, phase);
    return hipGetLastError();
}

}  // namespace psx_lbm_fast
